"""The rest of the backend interface on the device: single tensor blocks (compute_Dij and
the block functions), compute_r, one Ermak-McCammon move (compute_hydro_ermak), the IG-T
phases and load_forces, through the C ABI against the oracle and the golden vectors.

Tolerances: blocks <= 1e-13 relative to the largest entry (FP64, same expression order up to
FMA contraction); moves and Rs <= 1e-10 relative (the device log/sqrt of the Gaussian stream
agree to 2e-15 with glibc's)."""
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
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300)


def test_blocks_free_space(api):
    S, po = api
    sysm = wl.chain(16, seed=21)
    P = wl.parse_params(wl.chain_params())
    D = po.compute_D(sysm.coord, po.params_from(P))
    s = S.BDSystem(sysm, P, precision=S.FP64)
    print("blocks free space", flush=True)
    for i, j in ((3, 7), (0, 15), (9, 9), (12, 2)):
        ref = D[3 * j:3 * j + 3, 3 * i:3 * i + 3]
        assert _rel(s.compute_Dij(i, j, 0), ref) < 1e-13
        if i != j:
            assert _rel(s.compute_Dij(i, j, 1), ref) < 1e-13
    with pytest.raises(S.BackendError):
        s.compute_Dij(1, 2, 3)          # no Ewald tensor in free space
    s.close()


@pytest.mark.parametrize("case,kind", [("beenakker12", 3), ("smith12", 2)])
def test_blocks_ewald_vs_reference(api, golden, case, kind):
    S, _ = api
    g, sysm, P, p = golden(case)
    coord = g["params.coord"].reshape(-1, 4)
    Dref = g["D.D"]
    s = S.BDSystem(sysm, P, precision=S.FP64, coord=coord)
    print(f"blocks {case}", flush=True)
    for i, j in ((0, 5), (4, 4), (11, 2)):
        ref = Dref[3 * j:3 * j + 3, 3 * i:3 * i + 3]
        assert _rel(s.compute_Dij(i, j, 0), ref) < 1e-12 * np.max(np.abs(Dref)) / np.max(np.abs(ref))
        assert np.array_equal(s.compute_Dij(i, j, kind), s.compute_Dij(i, j, 0))
    with pytest.raises(S.BackendError):
        s.compute_Dij(0, 1, 5 - kind)   # the other method's block
    s.close()


def test_compute_r_is_L_times_the_stream(api):
    S, po = api
    sysm = wl.chain(20, seed=22)
    P = wl.parse_params(wl.chain_params(rand_seed=555))
    p = po.params_from(P)
    n = 3 * sysm.N
    L = po.cholesky(po.compute_D(sysm.coord, p))
    xi, _ = po.gauss(555, n, 2)
    s = S.BDSystem(sysm, P, precision=S.FP64)
    print("compute_r", flush=True)
    for k in range(2):
        ref = L @ (np.sqrt(2 * P.dt) * xi[k])
        assert _rel(s.compute_r(), ref) < 1e-10
    s.close()


@pytest.mark.parametrize("hydro", ["cholesky", "geyer", "none"])
def test_ermak_move_is_one_attempt(api, hydro):
    """compute_hydro_ermak: the move of the oracle's first step, the positions kept, the
    stream advanced by one draw; a second move equals the oracle's step from the start with
    the next draw."""
    S, po = api
    sysm = wl.chain(24, seed=23)
    P = wl.parse_params(wl.chain_params(hydro=hydro, rand_seed=808))
    p = po.params_from(P)
    s = S.BDSystem(sysm, P, precision=S.FP64)
    print(f"ermak move {hydro}", flush=True)
    out = s.ermak_move()
    ref = po.Sim(sysm, p, seed=808)
    ref.run(1)
    assert np.max(np.abs(out - ref.coords())) < 1e-10
    assert np.array_equal(s.coords(), sysm.coord)
    Rf, Rs = s.last_displacement()
    Rf_ref, Rs_ref = ref.last()
    assert _rel(Rs, Rs_ref) < 1e-10
    assert s.rng_state() == ref.rng_state()
    s.close()


def test_igt_phases_compose_the_step(api):
    """compute_hydro_IGT (phase 0) from the start equals one device IG-T step; the
    corrector alone (phase 1) redraws from the same start with the mean tensor kept."""
    S, po = api
    sysm = wl.chain(18, seed=24)
    P = wl.parse_params(wl.chain_params(algorithm="igt_const", rand_seed=99))
    a = S.BDSystem(sysm, P, precision=S.FP64)
    b = S.BDSystem(sysm, P, precision=S.FP64)
    print("igt phases", flush=True)
    out, res = a.igt_phase(0)
    b.compute_hydro(1)
    assert res == 0
    assert np.max(np.abs(out - b.coords())) < 1e-12
    # the corrector again from the start: a different draw, the same mean deterministic move
    a.set_coords(sysm.coord)
    out2, res2 = a.igt_phase(1)
    assert res2 == 0
    assert np.max(np.abs(out2 - out)) > 1e-6
    a.close()
    b.close()


def test_set_forces_round_trip(api):
    S, _ = api
    sysm = wl.chain(10, seed=25)
    P = wl.parse_params(wl.chain_params())
    s = S.BDSystem(sysm, P, precision=S.FP64)
    F = np.random.default_rng(0).normal(size=30)
    s.set_forces(F)
    assert np.array_equal(s.forces(), F)
    s.close()
