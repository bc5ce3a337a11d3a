"""Reverse-mode derivative of the step on the device (bdh_step_vjp, bdh_step_vjp_params)
and the differentiable trajectories built on it (bd_box_amd.adjoint).

  one step vs the oracle adjoint (itself pinned by central differences): <= 1e-9 relative
  K = 25 steps, checkpoints every 10, vs the oracle's chain of step VJPs: <= 1e-8 relative
  the same gradient vs central differences of the device trajectory: <= 1e-5 relative
  parameter gradients of (mean Rg^2 - 400)^2 (LJ epsilon and sigma, bond k, charges) vs the
  oracle's FP64 gradient: <= 1e-8 relative; vs central differences of the device
  trajectory (h = 1e-4 relative, charges 1e-4 absolute): <= 1e-5 relative
  mixed precision (FP32 factor promoted, FP64 reverse step) vs FP64: one step <= 1e-5
  relative, a 25-step trajectory gradient <= 1e-4 relative
Systems: the chain recipe (LJ, Debye-Hueckel, harmonic bonds) at its hydrodynamic radius
1.9 A: for equal radii the RPY block is C1 at the contact distance 2a (f, f', g, g' agree on
both branches), so bonded beads near 3.8 A have a derivative.
"""
import numpy as np
import pytest

from bd_box_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from bd_box_amd import system as S
    from oracle import adjoint as ad
    from oracle import pyoracle as po
    return S, ad, po


def _chain(N, seed):
    sysm = wl.chain(N, seed=seed)
    P = wl.parse_params(wl.chain_params())
    return sysm, P


@pytest.mark.parametrize("N", [12, 45, 90])
def test_step_vjp_vs_oracle(api, N):
    S, ad, po = api
    sysm, P = _chain(N, seed=900 + N)
    p = po.params_from(P)
    rng = np.random.default_rng(N)
    xi = rng.normal(size=3 * N)
    g = rng.normal(size=3 * N)
    print(f"step vjp N={N}", flush=True)
    s = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=1)
    rbar = s.step_vjp(sysm.coord, xi, g)
    ref = ad.step_vjp(sysm, p, sysm.coord, xi, g)
    assert np.linalg.norm(rbar - ref) < 1e-9 * np.linalg.norm(ref)


def test_cached_factor_gives_the_same_vjp(api):
    """The factor saved when stepping from coord is the one a refactor at coord gives."""
    S, ad, po = api
    sysm, P = _chain(30, seed=5)
    rng = np.random.default_rng(30)
    xi = rng.normal(size=3 * sysm.N)
    g = rng.normal(size=3 * sysm.N)
    print("cached factor vjp", flush=True)
    s = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=1)
    ref = s.step_vjp(sysm.coord, xi, g)
    s.factor_cache(2)
    s.set_coords(sysm.coord)
    s.compute_hydro(1, xi)
    s.save_factor(1)
    assert np.array_equal(s.step_vjp(sysm.coord, xi, g, slot=1), ref)
    s.factor_cache(0)


def test_trajectory_gradient(api):
    import torch
    from bd_box_amd import adjoint as adj
    S, ad, po = api
    N, K = 20, 25
    sysm, P = _chain(N, seed=77)
    p = po.params_from(P)
    rng = np.random.default_rng(8)
    xi = rng.normal(size=(K, 3 * N))
    w = rng.normal(size=(N, 3))
    print("trajectory gradient", flush=True)
    s = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=1)
    x0 = torch.tensor(sysm.coord[:, :3].copy(), dtype=torch.float64, requires_grad=True)
    out = adj.trajectory(s, x0, xi, sysm.coord[:, 3], checkpoint=10)
    (out * torch.from_numpy(w)).sum().backward()
    grad = x0.grad.numpy().reshape(-1)
    # oracle: forward positions, then the step VJPs backwards
    cs = [np.array(sysm.coord, float)]
    for k in range(K):
        cs.append(ad.step(sysm, p, cs[-1], xi[k]))
    g = w.reshape(-1).copy()
    for k in range(K - 1, -1, -1):
        g = ad.step_vjp(sysm, p, cs[k], xi[k], g)
    assert np.max(np.abs(out.detach().numpy() - cs[-1][:, :3])) < 1e-9
    assert np.linalg.norm(grad - g) < 1e-8 * np.linalg.norm(g)
    # central differences of the device trajectory along a random direction
    v = rng.normal(size=(N, 3))
    h = 1e-6

    def f(t):
        x = torch.tensor(sysm.coord[:, :3] + t * v, dtype=torch.float64)
        return float((adj.trajectory(s, x, xi, sysm.coord[:, 3]) * torch.from_numpy(w)).sum())

    fd = (f(h) - f(-h)) / (2 * h)
    assert abs(grad @ v.reshape(-1) - fd) < 1e-5 * abs(fd)


def test_trajectory_refuses_other_steps(api):
    """A factor reused over several steps is not the step differentiated."""
    import torch
    from bd_box_amd import adjoint as adj
    S, ad, po = api
    sysm, P = _chain(10, seed=3)
    x0 = torch.tensor(sysm.coord[:, :3].copy(), dtype=torch.float64, requires_grad=True)
    xi = np.zeros((2, 30))
    for prec in (S.FP64, S.MIXED):
        s = S.BDSystem(sysm, P, precision=prec, refactor_every=10)
        with pytest.raises(ValueError, match="refactor_every"):
            adj.trajectory(s, x0, xi, sysm.coord[:, 3])
        s.close()


@pytest.mark.parametrize("prec", ["MIXED", "MIXED_SPLIT"])
def test_mixed_step_vjp_close_to_fp64(api, prec):
    """The mixed-precision reverse step (the FP32 factor promoted, FP64 solves) against the
    FP64 one at the same positions: the factors differ by the FP32 rounding (|LL^T - D| /
    |D| ~ 1e-7), so the cotangents agree to <= 1e-5 relative."""
    S, ad, po = api
    sysm, P = _chain(90, seed=990)
    rng = np.random.default_rng(90)
    xi = rng.normal(size=3 * sysm.N)
    g = rng.normal(size=3 * sysm.N)
    print(f"mixed step vjp {prec}", flush=True)
    a = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=1)
    b = S.BDSystem(sysm, P, precision=getattr(S, prec), refactor_every=1)
    ra = a.step_vjp(sysm.coord, xi, g)
    rb = b.step_vjp(sysm.coord, xi, g)
    assert np.linalg.norm(rb - ra) <= 1e-5 * np.linalg.norm(ra)
    # the cached FP32 factor gives the same cotangent as a refactor at coord
    b.factor_cache(1)
    b.set_coords(sysm.coord)
    b.compute_hydro(1, xi)
    b.save_factor(0)
    assert np.array_equal(b.step_vjp(sysm.coord, xi, g, slot=0), rb)
    b.factor_cache(0)


def test_mixed_trajectory_gradient_close_to_fp64(api):
    """K = 25 mixed-precision steps, checkpoints every 10: the gradient of a linear
    functional of the final positions against the FP64 trajectory's, <= 1e-4 relative."""
    import torch
    from bd_box_amd import adjoint as adj
    S, ad, po = api
    N, K = 64, 25
    sysm, P = _chain(N, seed=78)
    rng = np.random.default_rng(9)
    xi = rng.normal(size=(K, 3 * N))
    w = torch.from_numpy(rng.normal(size=(N, 3)))
    print("mixed trajectory gradient", flush=True)
    grads = []
    for prec in (S.FP64, S.MIXED):
        s = S.BDSystem(sysm, P, precision=prec, refactor_every=1)
        x0 = torch.tensor(sysm.coord[:, :3].copy(), dtype=torch.float64, requires_grad=True)
        (adj.trajectory(s, x0, xi, sysm.coord[:, 3], checkpoint=10) * w).sum().backward()
        grads.append(x0.grad.numpy().reshape(-1).copy())
        s.close()
    assert np.linalg.norm(grads[1] - grads[0]) <= 1e-4 * np.linalg.norm(grads[0])



def _jittered(sysm, seed, scale=0.2):
    c = np.array(sysm.coord, float)
    c[:, :3] += np.random.default_rng(seed).normal(scale=scale, size=(sysm.N, 3))
    return c


@pytest.mark.parametrize("N", [12, 64])
def test_step_vjp_params_vs_oracle(api, N):
    """Position and parameter cotangents of one step (bonds off their rest length) against
    the oracle: step_vjp and param_vjp with v = c D g."""
    S, ad, po = api
    sysm, P = _chain(N, seed=950 + N)
    p = po.params_from(P)
    rng = np.random.default_rng(N)
    xi = rng.normal(size=3 * N)
    g = rng.normal(size=3 * N)
    coord = _jittered(sysm, N)
    print(f"step vjp params N={N}", flush=True)
    s = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=1)
    rbar, eb, Rb, Qb, Hb = s.step_vjp_params(coord, xi, g)
    s.close()
    c = P.dt / (0.001986248532617 * P.T)
    D = po.compute_D(coord, p)
    ref = ad.param_vjp(sysm, p, coord, c * (D @ g))
    for got, want in zip((eb, Rb, Qb, Hb), ref):
        assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want)
    want = ad.step_vjp(sysm, p, coord, xi, g)
    assert np.linalg.norm(rbar - want) <= 1e-9 * np.linalg.norm(want)


def test_param_trajectory_vs_oracle_and_differences(api):
    """The gradient of (mean Rg^2 over the frames - 400)^2 after K = 12 steps (checkpoints
    every 5) with respect to the scalar LJ epsilon and sigma, the bond constant k and the
    per-bead charges: against the oracle's FP64 gradient and central differences of the
    device's own forward trajectory."""
    import torch
    from bd_box_amd import adjoint as adj
    S, ad, po = api
    N, K = 64, 12
    sysm, P = _chain(N, seed=4242)
    p = po.params_from(P)
    xi = np.random.default_rng(12).normal(size=(K, 3 * N))
    s = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=1)
    r0rmax = np.array(sysm.bond_parms, float)[:, :2]
    radii = sysm.coord[:, 3]
    eps0, sig0, k0 = 0.6, 4.0, 10.0
    print("param trajectory", flush=True)

    def run(eps, sig, k, Q, grad=False):
        t = lambda a: torch.tensor(a, dtype=torch.float64, requires_grad=grad)
        te, ts, tk, tq = t(eps), t(sig), t(k), t(np.asarray(Q, float))
        x0 = torch.tensor(sysm.coord[:, :3].copy(), dtype=torch.float64)
        frames = adj.param_trajectory(s, x0, te * torch.ones(N, dtype=torch.float64),
                                      ts / 2 * torch.ones(N, dtype=torch.float64), tq,
                                      tk * torch.ones(len(sysm.bonds), dtype=torch.float64),
                                      xi, radii, r0rmax, checkpoint=5)
        loss = adj.rg_loss(frames)
        if grad:
            loss.backward()
            return float(loss), frames, (te.grad.item(), ts.grad.item(), tk.grad.item(), tq.grad.numpy())
        return float(loss), frames, None

    Q0 = np.array(sysm.Q, float)
    loss, frames, (ge, gs, gk, gq) = run(eps0, sig0, k0, Q0, grad=True)
    ref_loss, (eb, Rb, Qb, Hb) = ad.trajectory_param_grad(sysm, p, xi)
    assert abs(loss - ref_loss) <= 1e-9 * abs(ref_loss)
    want = (np.sum(eb), 0.5 * np.sum(Rb), np.sum(Hb))
    for got, w in zip((ge, gs, gk), want):
        assert abs(got - w) <= 1e-8 * abs(w), (got, w)
    assert np.linalg.norm(gq - Qb) <= 1e-8 * np.linalg.norm(Qb)
    for name, x0_, got in (("eps", eps0, ge), ("sigma", sig0, gs), ("k", k0, gk)):
        h = 1e-4 * x0_
        args_p = dict(eps=eps0, sig=sig0, k=k0)
        args_m = dict(args_p)
        args_p[{"eps": "eps", "sigma": "sig", "k": "k"}[name]] = x0_ + h
        args_m[{"eps": "eps", "sigma": "sig", "k": "k"}[name]] = x0_ - h
        fd = (run(Q=Q0, **args_p)[0] - run(Q=Q0, **args_m)[0]) / (2 * h)
        assert abs(got - fd) <= 1e-5 * abs(fd), (name, got, fd)
    dirn = np.random.default_rng(2).normal(size=N)
    h = 1e-4
    fd = (run(eps0, sig0, k0, Q0 + h * dirn)[0] - run(eps0, sig0, k0, Q0 - h * dirn)[0]) / (2 * h)
    assert abs(gq @ dirn - fd) <= 1e-5 * abs(fd), (gq @ dirn, fd)
    s.close()


def test_step_vjp_refuses_overlap_checks(api):
    """With check_overlap the step taken may be a redraw; the device VJP refuses it."""
    S, ad, po = api
    sysm = wl.chain(10, seed=3)
    P = wl.parse_params(wl.chain_params(check_overlap="yes"))
    s = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=1)
    with pytest.raises(S.BackendError, match="check_overlap"):
        s.step_vjp(sysm.coord, np.zeros(30), np.ones(30))
    s.close()
