"""The refactor schedule and the N=3000 chain of the mixed-precision headline, on the GPU.

Reference for the composed step: comp_step_ermak / compute_hydro_ermak
(src/diff_alg/diff_comp.c:152-179, diff_equation.c:216-264), the noise of compute_r
(diff_rand.c:64-85) and dpotrf / dtrmv (src/myblas.c:45-118), composed on the host from
pinned pieces (tests/compose.py).  Tolerances, written in the asserts:
  FP64 path         per-step displacement <= 1e-10 relative, forces <= 1e-10
  mixed paths       per-step displacement <= 1e-5 relative, forces <= 1e-5,
                    ||L L^T - D||_F / ||D||_F <= 1e-6 against the host FP64 D
"""
import numpy as np
import pytest

from bd_box_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("R", [3, 10])
def test_schedule_chunks_cross_cycles(R):
    """Given Gaussians in calls of uneven length (chunks that start and end inside refactor
    cycles), and the device stream over the same calls: both follow the composed schedule."""
    from bd_box_amd import system as S
    from compose import compose_steps
    from oracle import pyoracle as po
    sysm = wl.chain(40, seed=50 + R)
    P = wl.parse_params(wl.chain_params(rand_seed=777))
    K = 25
    xi = np.random.default_rng(R).normal(size=(K, 3 * sysm.N))
    traj, _ = compose_steps(sysm, P, xi, R)
    x0 = sysm.coord[:, :3]
    chunks = [1, 4, 7, 2, 11]
    s = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=R)
    k = 0
    for m in chunks:
        s.compute_hydro(m, xi[k:k + m])
        k += m
        err = _rel(s.coords()[:, :3] - x0, traj[k - 1][:, :3] - x0)
        assert err < 1e-10, (k, err)
    # device stream: the oracle's Sim is pinned to the composition (test_oracle.py)
    ref = po.Sim(sysm, po.params_from(P), seed=P.rand_seed, refactor_every=R)
    d = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=R)
    k = 0
    for m in chunks:
        ref.run(m)
        d.compute_hydro(m)
        k += m
        assert d.rng_state() == ref.rng_state(), k
        err = _rel(d.coords()[:, :3] - x0, ref.coords()[:, :3] - x0)
        assert err < 1e-10, (k, err)
    s.close()
    d.close()


@pytest.fixture(scope="module")
def chain3000():
    """The headline chain (N = 3000, n = 9000), one refactor cycle of 10 steps with given
    Gaussians, composed on the host in FP64."""
    from compose import compose_steps
    sysm = wl.chain(3000, seed=1)
    P = wl.parse_params(wl.chain_params(rand_seed=20241))
    xi = np.random.default_rng(3000).normal(size=(10, 3 * sysm.N))
    traj, disp = compose_steps(sysm, P, xi, 10)
    return sysm, P, xi, traj, disp


@pytest.mark.parametrize("prec_name,tol", [("FP64", 1e-10), ("MIXED", 1e-5), ("MIXED_SPLIT", 1e-5)])
def test_chain3000_cycle(chain3000, prec_name, tol):
    """One 10-step refactor cycle of the N=3000 chain: the whole cycle in one call (one
    factorization, the 10-right-hand-side L.xi, graph replay) and the same steps one call
    at a time, per-step displacements against the host FP64 composition; forces at the
    device's own final positions against the oracle's."""
    from bd_box_amd import system as S
    from oracle import pyoracle as po
    sysm, P, xi, traj, disp = chain3000
    prec = getattr(S, prec_name)
    x0 = sysm.coord[:, :3]
    s = S.BDSystem(sysm, P, precision=prec, refactor_every=10)
    s.compute_hydro(10, xi)
    err = _rel(s.coords()[:, :3] - x0, traj[-1][:, :3] - x0)
    assert err < tol, ("whole cycle", err)
    F_dev = s.forces()
    F_ref, _, _ = po.forces(sysm, po.params_from(P), s.coords())
    assert _rel(F_dev, F_ref) < tol, ("forces", _rel(F_dev, F_ref))
    s.close()
    s1 = S.BDSystem(sysm, P, precision=prec, refactor_every=10)
    errs = []
    for k in range(10):
        s1.compute_hydro(1, xi[k])
        Rf, Rs = s1.last_displacement()
        errs.append(_rel(Rf + Rs, disp[k]))
    assert max(errs) < tol, errs
    s1.close()


@pytest.mark.parametrize("prec_name", ["MIXED", "MIXED_SPLIT"])
def test_chain3000_factor_residual(chain3000, prec_name):
    """||L L^T - D||_F / ||D||_F of the device factor (FP32 storage, FP64 diagonal blocks)
    against the FP64 tensor of the oracle at the same positions."""
    from bd_box_amd import system as S
    from oracle import pyoracle as po
    sysm, P, _, _, _ = chain3000
    s = S.BDSystem(sysm, P, precision=getattr(S, prec_name), refactor_every=10)
    L = s.factor()
    s.close()
    D = po.compute_D(sysm.coord, po.params_from(P))
    assert np.array_equal(L, np.tril(L))
    res = np.linalg.norm(L @ L.T - D) / np.linalg.norm(D)
    assert res < 1e-6, res
