"""GPU parity of the HIP backend against the FP64 oracle and the reference's golden vectors.

Every call goes through the C ABI (libbdbox_hip.so).  Tolerances:
  FP64 path   : D, forces, factor, products <= 1e-12 relative (FMA contraction and
                summation order are the only differences from the oracle)
  mixed path  : D entries <= 2e-6 relative to max|D| (FP32 arithmetic on FP64
                separations), ||L L^T - D|| / ||D|| <= 1e-6, displacement <= 1e-5
  Gaussian stream: the integer drand48 state is bit-identical; values agree to 2e-15
                (device log1p/log vs glibc log).
"""
import numpy as np
import pytest

from bd_box_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from bd_box_amd import system as S
    from oracle import pyoracle as po
    return S, po


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


# ----------------------------------------------------------------------------- MFMA layouts
@pytest.mark.parametrize("kind,M,K,N", [(0, 32, 2, 32), (1, 16, 4, 16), (2, 16, 4, 16),
                                        (3, 32, 16, 32), (4, 16, 32, 16)])
def test_mfma_layout(api, kind, M, K, N):
    S, _ = api
    rng = np.random.default_rng(kind)
    # small integers: exact in bf16, f32 and f64, so the product is exact
    A = rng.integers(-8, 9, size=(M, K)).astype(float)
    B = rng.integers(-8, 9, size=(K, N)).astype(float)
    C = S.probe_mfma(kind, A, B)
    assert np.array_equal(C.astype(float), A @ B), f"layout mismatch kind {kind}"


# ----------------------------------------------------------------------------- random stream
@pytest.mark.parametrize("seed,count,calls", [(1, 15, 4), (12345, 9000, 3), (77, 1, 3),
                                              (2024, 24577, 2)])
def test_gauss_stream(api, seed, count, calls):
    S, po = api
    ref, ref_state = po.gauss(seed, count, calls)
    st = S.seed_state(seed)
    for k in range(calls):
        g, st = S.gauss(st, count)
        np.testing.assert_allclose(g, ref[k], rtol=2e-15, atol=2e-15)
    assert st == ref_state


def test_gauss_stream_golden(api, golden):
    S, _ = api
    g = golden("gauss_seed12345")[0]["gauss.g"]
    st = S.seed_state(12345)
    for k in range(g.shape[0]):
        out, st = S.gauss(st, g.shape[1])
        np.testing.assert_allclose(out, g[k], rtol=2e-15, atol=2e-15)


# ----------------------------------------------------------------------------- tensor + forces
def _system(N, seed, **prm):
    sysm = wl.chain(N, seed=seed)
    P = wl.parse_params(wl.chain_params(**prm))
    return sysm, P


@pytest.mark.parametrize("case", ["chain16", "chain40", "overlap10"])
def test_compute_D_vs_reference_golden(api, golden, case):
    S, _ = api
    g, sysm, P, p = golden(case)
    Dref = g["D.D"]
    d64 = S.BDSystem(sysm, P, precision=S.FP64).compute_D()
    assert _rel(d64, Dref) < 1e-13
    d32 = S.BDSystem(sysm, P, precision=S.MIXED).compute_D()
    assert _rel(d32, Dref) < 2e-6


@pytest.mark.parametrize("N", [3, 64, 257])
def test_compute_D_vs_oracle(api, N):
    S, po = api
    sysm, P = _system(N, seed=N)
    Dref = po.compute_D(sysm.coord, po.params_from(P))
    assert _rel(S.BDSystem(sysm, P, precision=S.FP64).compute_D(), Dref) < 1e-13
    assert _rel(S.BDSystem(sysm, P, precision=S.MIXED).compute_D(), Dref) < 2e-6


@pytest.mark.parametrize("case", ["chain16", "chain40"])
def test_forces_vs_reference_golden(api, golden, case):
    S, _ = api
    g, sysm, P, p = golden(case)
    Fref, Eref = g["forces.F"], g["forces.E"]
    for prec, tol in ((S.FP64, 1e-12), (S.MIXED, 2e-5)):
        F, E, rep = S.BDSystem(sysm, P, precision=prec).compute_forces()
        assert _rel(F, Fref) < tol, prec
        assert abs(E[0] - Eref[6]) <= tol * max(1, abs(Eref[6]))
        assert abs(E[1] - Eref[7]) <= tol * max(1, abs(Eref[7]))
        assert rep == 0


def test_forces_scaled_bonded_pairs(api):
    S, po = api
    sysm = wl.chain(50, seed=8)
    sysm.bond_parms[:, 1] = 20.0
    P = wl.parse_params(wl.chain_params(bond_lj_scale=0.5, bond_c_scale=0.25))
    Fref, Eref, _ = po.forces(sysm, po.params_from(P))
    F, E, _ = S.BDSystem(sysm, P, precision=S.FP64).compute_forces()
    assert _rel(F, Fref) < 1e-12
    np.testing.assert_allclose(E, Eref, rtol=1e-11, atol=1e-12)


def test_forces_unscaled_bonds_interact(api):
    S, po = api
    sysm = wl.chain(40, seed=9)
    P = wl.parse_params(wl.chain_params(bond_lj_scale=1.0, bond_c_scale=1.0))
    Fref, _, _ = po.forces(sysm, po.params_from(P))
    F, _, _ = S.BDSystem(sysm, P, precision=S.FP64).compute_forces()
    assert _rel(F, Fref) < 1e-12


# ----------------------------------------------------------------------------- Cholesky + L.x
@pytest.mark.parametrize("N", [16, 100, 300, 700])
def test_factor_vs_oracle(api, N):
    S, po = api
    sysm, P = _system(N, seed=100 + N)
    D = po.compute_D(sysm.coord, po.params_from(P))
    Lref = po.cholesky(D)
    L64 = S.BDSystem(sysm, P, precision=S.FP64).factor()
    assert _rel(L64, Lref) < 1e-12
    assert np.all(np.triu(L64, 1) == 0)
    L32 = S.BDSystem(sysm, P, precision=S.MIXED).factor()
    res = np.linalg.norm(L32 @ L32.T - D) / np.linalg.norm(D)
    assert res < 1e-6, res
    # split-bf16 trailing updates: products to ~2^-17, FP32 accumulation
    Ls = S.BDSystem(sysm, P, precision=S.MIXED_SPLIT).factor()
    res = np.linalg.norm(Ls @ Ls.T - D) / np.linalg.norm(D)
    assert res < 5e-6, res


def test_factor_large_right_looking(api):
    """n_pad above 4096 takes the right-looking inner updates (smaller matrices update
    left-looking within the outer panel); LAPACK's factor of the oracle's D is the check."""
    S, po = api
    sysm, P = _system(1400, seed=1400)
    D = po.compute_D(sysm.coord, po.params_from(P))
    Lref = np.linalg.cholesky(D)
    print("factor N=1400", flush=True)
    L64 = S.BDSystem(sysm, P, precision=S.FP64).factor()
    assert _rel(L64, Lref) < 1e-12
    Ls = S.BDSystem(sysm, P, precision=S.MIXED_SPLIT).factor()
    v = np.random.default_rng(0).normal(size=D.shape[0])
    res = np.linalg.norm(Ls @ (Ls.T @ v) - D @ v) / np.linalg.norm(D @ v)
    assert res < 5e-6, res


@pytest.mark.parametrize("n", [1000, 1100, 1200, 2000, 9000])
def test_dpotrf_decaying_vs_lapack(api, n):
    """A dense SPD matrix with entries 0.9 / (1 + 0.05 d^2) off the diagonal (d = |i - j|) and
    2 on it (its symbol is at least 1.1): every entry of the device factor against LAPACK's.
    Off-diagonal entries far below the diagonal's size expose a missed or stale block update
    that a norm-wise residual can hide.  n_pad = 1024, 1152, 1280 put the one-launch-per-256-
    columns schedule at its smallest size and end it on a 128-column outer panel."""
    S, _ = api
    d = np.abs(np.subtract.outer(np.arange(n), np.arange(n))).astype(float)
    M = np.where(d == 0, 2.0, 0.9 / (1.0 + 0.05 * d * d))
    Lref = np.linalg.cholesky(M)
    print(f"dpotrf decaying n={n}", flush=True)
    for prec, tol in ((S.FP64, 1e-13), (S.MIXED, 2e-6), (S.MIXED_SPLIT, 1e-5)):
        A, info = S.dpotrf(M, "L", precision=prec)
        assert info == 0
        err = np.max(np.abs(np.tril(A) - Lref))
        assert err < tol, (prec, err)


def test_factor_vs_reference_golden(api, golden):
    S, _ = api
    for case in ("chain16", "overlap10"):
        g, sysm, P, p = golden(case)
        L = S.BDSystem(sysm, P, precision=S.FP64).factor()
        assert _rel(L, g["chol.L"]) < 1e-12


@pytest.mark.parametrize("N", [16, 300])
def test_noise_product(api, N):
    S, po = api
    sysm, P = _system(N, seed=200 + N)
    D = po.compute_D(sysm.coord, po.params_from(P))
    Lref = po.cholesky(D)
    x = np.random.default_rng(N).normal(size=3 * N)
    for prec, tol in ((S.FP64, 1e-12), (S.MIXED, 1e-5), (S.MIXED_SPLIT, 1e-5)):
        s = S.BDSystem(sysm, P, precision=prec)
        s.factor()
        y = s.noise_product(x)
        assert np.linalg.norm(y - Lref @ x) / np.linalg.norm(Lref @ x) < tol, prec


def test_host_dpotrf_dtrmv(api):
    S, po = api
    sysm, P = _system(60, seed=5)
    D = po.compute_D(sysm.coord, po.params_from(P))
    Lref = po.cholesky(D)
    x = np.random.default_rng(1).normal(size=D.shape[0])
    A, info = S.dpotrf(D, "U")
    assert info == 0
    assert _rel(np.triu(A), Lref.T) < 1e-12            # U = L^T, LAPACK "U" convention
    assert _rel(np.tril(A, -1), np.tril(D, -1)) == 0   # strict lower part untouched
    y = S.dtrmv(A, x, "U")
    assert np.linalg.norm(y - Lref @ x) / np.linalg.norm(Lref @ x) < 1e-12
    Bm, info = S.dpotrf(D, "L")
    assert info == 0 and _rel(np.tril(Bm), Lref) < 1e-12
    # a non-SPD matrix reports the failing pivot (LAPACK INFO > 0)
    bad = D.copy()
    bad[5, 5] = -1.0
    _, info = S.dpotrf(bad, "L")
    assert info == 6
    # a pivot in a later diagonal block (factored inside the trailing-update kernel)
    bad = D.copy()
    bad[150, 150] = -1.0
    for prec in (S.FP64, S.MIXED, S.MIXED_SPLIT):
        _, info = S.dpotrf(bad, "L", precision=prec)
        assert info == 151, (prec, info)


@pytest.mark.parametrize("N", [20, 150])
def test_compute_Rf(api, N):
    S, po = api
    sysm, P = _system(N, seed=300 + N)
    p = po.params_from(P)
    F, _, _ = po.forces(sysm, p)
    D = po.compute_D(sysm.coord, p)
    ref = (P.dt / (0.001986248532617 * P.T)) * (D @ F)
    for prec, tol in ((S.FP64, 1e-12), (S.MIXED, 1e-5)):
        Rf = S.BDSystem(sysm, P, precision=prec).compute_Rf()
        assert np.linalg.norm(Rf - ref) / np.linalg.norm(ref) < tol, prec


# ----------------------------------------------------------------------------- full steps
@pytest.mark.parametrize("N,refactor,steps", [(16, 1, 5), (70, 3, 7)])
def test_steps_vs_oracle_with_given_noise(api, N, refactor, steps):
    S, po = api
    sysm, P = _system(N, seed=400 + N)
    xi = np.random.default_rng(N).normal(size=(steps, 3 * N))
    ref = po.Sim(sysm, po.params_from(P), seed=1, refactor_every=refactor)
    ref.run(steps, xi)
    for prec, tol in ((S.FP64, 1e-10), (S.MIXED, 1e-5)):
        s = S.BDSystem(sysm, P, precision=prec, refactor_every=refactor)
        s.compute_hydro(steps, xi)
        d_gpu = s.coords()[:, :3] - sysm.coord[:, :3]
        d_ref = ref.coords()[:, :3] - sysm.coord[:, :3]
        err = np.linalg.norm(d_gpu - d_ref) / np.linalg.norm(d_ref)
        assert err < tol, (prec, err)
        assert _rel(s.forces(), ref.forces()) < (1e-9 if prec == S.FP64 else 1e-4)


def test_trajectory_vs_reference_golden(api, golden):
    """Device drand48 stream + FP64 step against the reference's own trajectory."""
    S, _ = api
    for case, steps in (("chain16", 6), ("chain40", 4)):
        g, sysm, P, p = golden(case)
        traj = g["traj.coord"].reshape(steps, -1, 4)
        s = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=1)
        for k in range(steps):
            s.compute_hydro(1)
            assert np.max(np.abs(s.coords() - traj[k])) < 1e-10, (case, k)


def test_first_step_parts_vs_reference_golden(api, golden):
    S, _ = api
    g, sysm, P, p = golden("chain16")
    for prec, tol in ((S.FP64, 1e-11), (S.MIXED, 1e-5)):
        s = S.BDSystem(sysm, P, precision=prec)
        s.compute_hydro(1)
        Rf, Rs = s.last_displacement()
        assert _rel(Rs, g["step1.Rs"]) < tol
        assert _rel(Rf, g["step1.Rf"]) < max(tol, 1e-4 if prec == S.MIXED else 0)


def test_device_stream_matches_oracle_stream(api):
    S, po = api
    sysm, P = _system(33, seed=7, rand_seed=999)
    ref = po.Sim(sysm, po.params_from(P), seed=999)
    ref.run(3)
    s = S.BDSystem(sysm, P, precision=S.FP64)
    s.compute_hydro(3)
    assert s.rng_state() == ref.rng_state()
    assert np.max(np.abs(s.coords() - ref.coords())) < 1e-10


@pytest.mark.parametrize("N,refactor,steps", [(33, 5, 12), (40, 4, 9)])
def test_device_stream_batched_cycles(api, N, refactor, steps):
    """Noise drawn for a whole refactor cycle in one launch (odd and even 3N) and L xi of
    the cycle in one pass over L: same stream state and trajectory as step-by-step."""
    S, po = api
    sysm, P = _system(N, seed=8, rand_seed=4242)
    ref = po.Sim(sysm, po.params_from(P), seed=4242, refactor_every=refactor)
    ref.run(steps)
    s = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=refactor)
    s.compute_hydro(steps)
    assert s.rng_state() == ref.rng_state()
    assert np.max(np.abs(s.coords() - ref.coords())) < 1e-10
    # the same steps one call at a time take the unbatched path
    s1 = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=refactor)
    for _ in range(steps):
        s1.compute_hydro(1)
    assert s1.rng_state() == ref.rng_state()
    assert np.max(np.abs(s1.coords() - s.coords())) < 1e-11


def test_no_hydro_path(api):
    S, po = api
    sysm, P = _system(25, seed=3, hydro="none")
    xi = np.random.default_rng(2).normal(size=(2, 75))
    ref = po.Sim(sysm, po.params_from(P), seed=1)
    ref.run(2, xi)
    s = S.BDSystem(sysm, P, precision=S.FP64)
    s.compute_hydro(2, xi)
    assert np.max(np.abs(s.coords() - ref.coords())) < 1e-11


# ----------------------------------------------------------------------------- TEA, Ewald
def _bk_coord(golden):
    return golden("beenakker12")[0]["params.coord"].reshape(-1, 4)


@pytest.mark.parametrize("case", ["beenakker12", "beenakker12_chol"])
def test_beenakker_D_vs_reference_golden(api, golden, case):
    S, _ = api
    g, sysm, P, p = golden(case)
    coord = _bk_coord(golden)
    Dref = g["D.D"]
    assert _rel(S.BDSystem(sysm, P, precision=S.FP64, coord=coord).compute_D(), Dref) < 1e-12
    assert _rel(S.BDSystem(sysm, P, precision=S.MIXED, coord=coord).compute_D(), Dref) < 1e-5


@pytest.mark.parametrize("case", ["beenakker12", "beenakker12_chol", "geyer24"])
def test_first_step_tea_ewald_vs_reference_golden(api, golden, case):
    """First step with the device stream: Rf and Rs against the reference's own
    (TEA free space, TEA + Beenakker, Cholesky + Beenakker)."""
    S, _ = api
    g, sysm, P, p = golden(case)
    coord = _bk_coord(golden) if case.startswith("beenakker") else None
    for prec, tol in ((S.FP64, 1e-10), (S.MIXED, 2e-5)):
        s = S.BDSystem(sysm, P, precision=prec, coord=coord)
        s.compute_hydro(1)
        Rf, Rs = s.last_displacement()
        assert _rel(Rs, g["step1.Rs"]) < tol, (case, prec, _rel(Rs, g["step1.Rs"]))
        if np.abs(g["step1.Rf"]).max() > 0:
            assert _rel(Rf, g["step1.Rf"]) < max(tol, 1e-4 if prec == S.MIXED else 0), (case, prec)


def test_geyer_trajectory_vs_reference_golden(api, golden):
    S, _ = api
    g, sysm, P, p = golden("geyer24")
    traj = g["traj.coord"].reshape(3, -1, 4)
    s = S.BDSystem(sysm, P, precision=S.FP64)
    for k in range(3):
        s.compute_hydro(1)
        assert np.max(np.abs(s.coords() - traj[k])) < 1e-10, k


@pytest.mark.parametrize("N", [50, 130])
def test_tea_vs_oracle_given_noise(api, N):
    S, po = api
    sysm, P = _system(N, seed=500 + N, hydro="geyer")
    xi = np.random.default_rng(N).normal(size=(3, 3 * N))
    ref = po.Sim(sysm, po.params_from(P), seed=1)
    ref.run(3, xi)
    for prec, tol in ((S.FP64, 1e-10), (S.MIXED, 1e-5)):
        s = S.BDSystem(sysm, P, precision=prec)
        s.compute_hydro(3, xi)
        d_gpu = s.coords()[:, :3] - sysm.coord[:, :3]
        d_ref = ref.coords()[:, :3] - sysm.coord[:, :3]
        assert np.linalg.norm(d_gpu - d_ref) / np.linalg.norm(d_ref) < tol, prec


# ----------------------------------------------------------------------------- Smith-Ewald
@pytest.mark.parametrize("case", ["smith12", "smith_ov8"])
def test_smith_D_vs_reference_golden(api, golden, case):
    """Smith-Ewald tensor (diff_tensor.c:102-263) against the reference's own D: FP64 to
    1e-12 relative, the mixed path (FP32 real-space images) to 1e-5."""
    S, _ = api
    g, sysm, P, p = golden(case)
    assert P.ewald_method == 0
    coord = g["params.coord"].reshape(-1, 4)
    print(f"smith D {case}", flush=True)
    Dref = g["D.D"]
    assert _rel(S.BDSystem(sysm, P, precision=S.FP64, coord=coord).compute_D(), Dref) < 1e-12
    assert _rel(S.BDSystem(sysm, P, precision=S.MIXED, coord=coord).compute_D(), Dref) < 1e-5


@pytest.mark.parametrize("case", ["smith12", "smith12_geyer"])
def test_smith_steps_vs_reference_golden(api, golden, case):
    """First step and a 3-step trajectory with the device stream (Cholesky and TEA noise
    over the Smith-Ewald tensor) against the reference's own."""
    S, _ = api
    g, sysm, P, p = golden(case)
    coord = golden("smith12")[0]["params.coord"].reshape(-1, 4)
    print(f"smith steps {case}", flush=True)
    for prec, tol in ((S.FP64, 1e-10), (S.MIXED, 2e-5)):
        s = S.BDSystem(sysm, P, precision=prec, coord=coord)
        s.compute_hydro(1)
        Rf, Rs = s.last_displacement()
        assert _rel(Rs, g["step1.Rs"]) < tol, (case, prec, _rel(Rs, g["step1.Rs"]))
    s = S.BDSystem(sysm, P, precision=S.FP64, coord=coord)
    traj = g["traj.coord"].reshape(3, -1, 4)
    for k in range(3):
        s.compute_hydro(1)
        assert np.max(np.abs(s.coords() - traj[k])) < 1e-9, (case, k)
