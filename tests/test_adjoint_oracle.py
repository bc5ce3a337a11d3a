"""The FP64 oracle adjoint (oracle/adjoint.py) against central finite differences of the
oracle's own forward pass, piece by piece and for the whole step.  Tolerance 1e-6 relative
(central differences with h = 1e-5 A have O(h^2) ~ 1e-10 truncation; the rest is rounding)."""
import numpy as np
import pytest

from bd_box_amd import workloads as wl


@pytest.fixture(scope="module")
def setup():
    from oracle import adjoint as ad
    from oracle import pyoracle as po
    sysm = wl.chain(12, seed=31)
    # hydrodynamic radius 1.6 A: bonded beads (3.8 A apart) stay clear of the contact
    # distance where the RPY block switches form (its derivative jumps there)
    sysm.coord[:, 3] = 1.6
    P = wl.parse_params(wl.chain_params())
    return ad, po, sysm, P, po.params_from(P)


def _fd(fun, x, v, h=1e-5):
    return (fun(x + h * v) - fun(x - h * v)) / (2 * h)


def _perturb(coord, v):
    c = np.array(coord, float).reshape(-1, 4).copy()
    c[:, :3] += v.reshape(-1, 3)
    return c


def test_forward_step_matches_oracle_integrator(setup):
    ad, po, sysm, P, p = setup
    xi = np.random.default_rng(1).normal(size=3 * sysm.N)
    sim = po.Sim(sysm, p, seed=1, refactor_every=1)
    sim.run(1, xi)
    assert np.max(np.abs(ad.step(sysm, p, sysm.coord, xi) - sim.coords())) < 1e-11


def test_rpy_contraction_gradient(setup):
    ad, po, sysm, P, p = setup
    rng = np.random.default_rng(2)
    n = 3 * sysm.N
    U = rng.normal(size=(n, n))
    v = rng.normal(size=n)
    g = ad.rpy_contract_grad(sysm.coord, U, p)
    fd = _fd(lambda t: np.sum(U * po.compute_D(_perturb(sysm.coord, t * v), p)), 0.0, 1.0)
    assert abs(g @ v - fd) < 1e-6 * abs(fd)


def test_rpy_contraction_gradient_overlapping_beads(setup):
    """Both branches of the RPY block (r < a_i + a_j uses the overlap form)."""
    ad, po, sysm, P, p = setup
    c = np.array(sysm.coord, float)
    c[:, 3] = 2.5                       # radii 2.5 A, bonds 3.8 A: bonded pairs overlap
    rng = np.random.default_rng(3)
    n = 3 * sysm.N
    U = rng.normal(size=(n, n))
    v = rng.normal(size=n)
    g = ad.rpy_contract_grad(c, U, p)
    fd = _fd(lambda t: np.sum(U * po.compute_D(_perturb(c, t * v), p)), 0.0, 1.0)
    assert abs(g @ v - fd) < 1e-6 * abs(fd)


def test_hessian_vector_product(setup):
    ad, po, sysm, P, p = setup
    rng = np.random.default_rng(4)
    y = rng.normal(size=3 * sysm.N)
    Hy = ad.hessian_vec(sysm, p, sysm.coord, y)
    F = lambda t: po.forces(sysm, p, coord=_perturb(sysm.coord, t * y))[0]
    fd = -_fd(F, 0.0, 1.0)
    assert np.linalg.norm(Hy - fd) < 1e-6 * np.linalg.norm(fd)


def test_cholesky_vjp(setup):
    ad, po, sysm, P, p = setup
    rng = np.random.default_rng(5)
    D = po.compute_D(sysm.coord, p)
    n = D.shape[0]
    Lbar = np.tril(rng.normal(size=(n, n)))
    S = rng.normal(size=(n, n))
    S = S + S.T
    U = ad.chol_vjp(po.cholesky(D), Lbar)
    fd = _fd(lambda t: np.sum(Lbar * np.linalg.cholesky(D + t * S)), 0.0, 1.0, h=1e-7)
    assert abs(np.sum(U * S) - fd) < 1e-6 * abs(fd)


def test_step_vjp(setup):
    ad, po, sysm, P, p = setup
    rng = np.random.default_rng(6)
    n = 3 * sysm.N
    xi = rng.normal(size=n)
    g = rng.normal(size=n)
    v = rng.normal(size=n)
    rbar = ad.step_vjp(sysm, p, sysm.coord, xi, g)
    f = lambda t: np.sum(g.reshape(-1, 3) * ad.step(sysm, p, _perturb(sysm.coord, t * v), xi)[:, :3])
    fd = _fd(f, 0.0, 1.0)
    assert abs(rbar @ v - fd) < 1e-6 * abs(fd)


def test_trajectory_refuses_steps_it_does_not_differentiate(golden):
    """With check_overlap a rejected step is redrawn from the stream (ermak_var also halves
    dt): the oracle's step with given Gaussians then departs from the step taken, so the
    differentiated configuration excludes it (bdh_step_vjp refuses it the same way)."""
    from oracle import pyoracle as po
    from bd_box_amd.adjoint import check_differentiable
    from bd_box_amd.system import make_config
    g, sysm, P, p = golden("retry_const")
    assert P.check_overlap == 1
    n = 3 * sysm.N
    xi = np.random.default_rng(5).normal(size=(8, n))
    taken = po.Sim(sysm, p, seed=P.rand_seed, algorithm=P.algorithm, move_attempts=P.move_attempts)
    taken.run(8, xi)
    assert taken.rejections() > 0
    p.check_overlap = 0
    given = po.Sim(sysm, p, seed=P.rand_seed)
    given.run(8, xi)
    assert np.max(np.abs(taken.coords() - given.coords())) > 1e-6
    with pytest.raises(ValueError, match="check_overlap"):
        check_differentiable(make_config(P, precision=0, refactor_every=1))
    P.check_overlap = 0
    P.algorithm = 1
    with pytest.raises(ValueError, match="ermak_const"):
        check_differentiable(make_config(P, precision=0, refactor_every=1))
    P.algorithm = 0
    check_differentiable(make_config(P, precision=0, refactor_every=1))
    check_differentiable(make_config(P, precision=1, refactor_every=1))
    with pytest.raises(ValueError, match="refactor_every"):
        check_differentiable(make_config(P, precision=1, refactor_every=10))


@pytest.fixture(scope="module")
def chain19():
    """The chain recipe at its own hydrodynamic radius (1.9 A): for equal radii the RPY
    block is C1 at contact (f, f', g, g' agree on both branches at r = 2a), so bonded beads
    at 3.8 A need no special radius."""
    from oracle import adjoint as ad
    from oracle import pyoracle as po
    sysm = wl.chain(12, seed=77)
    P = wl.parse_params(wl.chain_params())
    return ad, po, sysm, P, po.params_from(P)


def test_rpy_block_is_c1_at_contact(chain19):
    ad, po, sysm, P, p = chain19
    dii, dij = po.prefactors(p)
    a = 1.9
    lo = ad.rpy_fg(np.array([2 * a * (1 - 1e-12)]), a, a, dii, dij)
    hi = ad.rpy_fg(np.array([2 * a * (1 + 1e-12)]), a, a, dii, dij)
    for x, y in zip(lo, hi):
        assert abs(x[0] - y[0]) <= 1e-9 * abs(y[0]), (x, y)


def _theta(sysm):
    return dict(eps=np.array(sysm.LJ, float)[:, 1].copy(), R=np.array(sysm.LJ, float)[:, 0].copy(),
                Q=np.array(sysm.Q, float).copy(), H=np.array(sysm.bond_parms, float)[:, 2].copy())


def test_param_vjp_matches_force_differences(chain19):
    """v^T dF/dtheta against central differences of the oracle's forces (h = 1e-6
    relative; Q absolute), along random directions of every parameter group."""
    ad, po, sysm, P, p = chain19
    rng = np.random.default_rng(3)
    v = rng.normal(size=3 * sysm.N)
    th = _theta(sysm)
    # bonds off their rest length, so the bond term has a derivative in H
    coord = np.array(sysm.coord, float)
    coord[:, :3] += rng.normal(scale=0.2, size=(sysm.N, 3))
    got = dict(zip(("eps", "R", "Q", "H"), ad.param_vjp(sysm, p, coord, v)))
    for name in ("eps", "R", "Q", "H"):
        dirn = rng.normal(size=th[name].shape)
        h = 1e-6 * (1.0 if name == "Q" else np.abs(th[name]).max())
        Fp, _, _ = po.forces(ad.with_params(sysm, **{name: th[name] + h * dirn}), p, coord)
        Fm, _, _ = po.forces(ad.with_params(sysm, **{name: th[name] - h * dirn}), p, coord)
        fd = v @ (Fp - Fm) / (2 * h)
        assert abs(got[name] @ dirn - fd) <= 1e-6 * max(abs(fd), 1e-12), (name, got[name] @ dirn, fd)


def test_param_vjp_discriminates(chain19):
    """A dropped factor in the LJ epsilon rule (eps_kj taken as eps_k) is told apart."""
    ad, po, sysm, P, p = chain19
    v = np.random.default_rng(4).normal(size=3 * sysm.N)
    eps_bar = ad.param_vjp(sysm, p, sysm.coord, v)[0]
    th = _theta(sysm)
    h = 1e-6
    dirn = np.zeros(sysm.N)
    dirn[3] = 1.0
    Fp, _, _ = po.forces(ad.with_params(sysm, eps=th["eps"] + h * dirn), p)
    Fm, _, _ = po.forces(ad.with_params(sysm, eps=th["eps"] - h * dirn), p)
    fd = v @ (Fp - Fm) / (2 * h)
    assert abs(eps_bar[3] - fd) <= 1e-6 * abs(fd)
    assert abs(2 * eps_bar[3] - fd) > 1e-3 * abs(fd)


def test_trajectory_param_grad_vs_central_differences(chain19):
    """Gradient of (mean Rg^2 over the frames - 400)^2 after K = 5 steps with given
    Gaussians, for the scalar LJ epsilon and sigma, the bond constant and the charges,
    against central differences of the oracle's forward trajectory (h = 1e-4 relative;
    charges: 1e-4 absolute along a random direction)."""
    ad, po, sysm, P, p = chain19
    K = 5
    xi = np.random.default_rng(9).normal(size=(K, 3 * sysm.N))
    loss, (eb, Rb, Qb, Hb) = ad.trajectory_param_grad(sysm, p, xi)
    th = _theta(sysm)

    def L(**kw):
        frames = ad.trajectory(ad.with_params(sysm, **kw), p, xi)
        return ad.rg_loss([f[:, :3] for f in frames])[0]

    eps0, sig0, k0 = th["eps"][0], 2 * th["R"][0], th["H"][0]
    cases = [
        ("eps", eps0, np.sum(eb), lambda x: dict(eps=np.full(sysm.N, x))),
        ("sigma", sig0, 0.5 * np.sum(Rb), lambda x: dict(R=np.full(sysm.N, x / 2))),
        ("k", k0, np.sum(Hb), lambda x: dict(H=np.full(len(sysm.bonds), x))),
    ]
    for name, x0, grad, mk in cases:
        h = 1e-4 * x0
        fd = (L(**mk(x0 + h)) - L(**mk(x0 - h))) / (2 * h)
        assert abs(grad - fd) <= 1e-5 * abs(fd), (name, grad, fd)
    dirn = np.random.default_rng(1).normal(size=sysm.N)
    h = 1e-4
    fd = (L(Q=th["Q"] + h * dirn) - L(Q=th["Q"] - h * dirn)) / (2 * h)
    assert abs(Qb @ dirn - fd) <= 1e-5 * abs(fd), (Qb @ dirn, fd)
