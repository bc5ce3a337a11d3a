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
