"""IG-T predictor-corrector on the device (comp_step_IGT_const / _var, diff_comp.c:76-144;
compute_hydro_IGT and its second phase, diff_equation.c:288-374), with Cholesky, diagonal
and TEA noise.

Each step factors D(r0) and the mean tensor (D(r0) + D(r*)) / 2 and draws two Gaussian
vectors (more after rejections); the device must reproduce the reference's trajectory, its
per-step curr_dt and its stream.

  positions (FP64) : <= 1e-10 absolute against the reference's own trajectory
  dt               : identical
  mixed precision  : per-step positions within 1e-5 relative of the FP64 device run
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [("igt_const", 6), ("igt_var_none", 6), ("igt_const_retry", 8),
         ("igt_var_retry", 8), ("igt_var_retry_none", 8)]


@pytest.fixture(scope="module")
def api():
    from bd_box_amd import system as S
    from oracle import pyoracle as po
    return S, po


@pytest.mark.parametrize("case,steps", CASES)
def test_igt_matches_reference(api, golden, case, steps):
    S, po = api
    g, sysm, P, p = golden(case)
    print(f"igt trajectory {case}", flush=True)
    s = S.BDSystem(sysm, P, precision=S.FP64)
    traj = g["traj.coord"].reshape(steps, -1, 4)
    for k in range(steps):
        s.compute_hydro(1)
        assert s.last_dt() == g["traj.dt"][k], (case, k)
        err = np.max(np.abs(s.coords()[:, :3] - traj[k][:, :3]))
        assert err < 1e-10, (case, k, err)
    sim = po.Sim(sysm, p, seed=P.rand_seed, algorithm=P.algorithm,
                 move_attempts=P.move_attempts)
    sim.run(steps)
    assert s.rng_state() == sim.rng_state()
    Rf, _ = s.last_displacement()
    Rf_o, _ = sim.last()
    assert np.max(np.abs(Rf - Rf_o)) <= 1e-10 * max(np.max(np.abs(Rf_o)), 1e-300)


@pytest.mark.parametrize("prec", ["MIXED", "MIXED_SPLIT"])
def test_igt_mixed_close_to_fp64(api, golden, prec):
    S, po = api
    g, sysm, P, p = golden("igt_const")
    print(f"igt mixed {prec}", flush=True)
    a = S.BDSystem(sysm, P, precision=S.FP64)
    b = S.BDSystem(sysm, P, precision=getattr(S, prec))
    for k in range(6):
        b.set_coords(a.coords())
        b.set_rng_state(a.rng_state())
        c0 = a.coords()[:, :3].copy()
        a.compute_hydro(1)
        b.compute_hydro(1)
        da = a.coords()[:, :3] - c0
        db = b.coords()[:, :3] - c0
        assert np.linalg.norm(db - da) <= 1e-5 * np.linalg.norm(da), k


TEA_CASES = [("igt_geyer", 6), ("igt_geyer_retry", 8), ("igt_var_geyer_retry", 8),
             ("igt_geyer_beenakker", 3)]


def _start(golden, case):
    # the periodic spheres start from the reference's centred table
    return golden("beenakker12")[0]["params.coord"].reshape(-1, 4) if "beenakker" in case else None


@pytest.mark.parametrize("case,steps", TEA_CASES)
def test_igt_tea_matches_reference(api, golden, case, steps):
    """IG-T with TEA noise: the predictor's TEA sums over D(r0), the corrector's over the mean
    tensor (D(r0) + D(r*)) / 2, blocks re-evaluated at both position sets."""
    S, po = api
    g, sysm, P, p = golden(case)
    print(f"igt tea trajectory {case}", flush=True)
    s = S.BDSystem(sysm, P, precision=S.FP64, coord=_start(golden, case))
    traj = g["traj.coord"].reshape(steps, -1, 4)
    for k in range(steps):
        s.compute_hydro(1)
        assert s.last_dt() == g["traj.dt"][k], (case, k)
        err = np.max(np.abs(s.coords()[:, :3] - traj[k][:, :3]))
        assert err < 1e-10, (case, k, err)
    sim = po.Sim(sysm, p, seed=P.rand_seed, algorithm=P.algorithm,
                 move_attempts=P.move_attempts, coord=_start(golden, case))
    sim.run(steps)
    assert s.rng_state() == sim.rng_state()


@pytest.mark.parametrize("case", ["igt_geyer", "igt_geyer_beenakker"])
def test_igt_tea_mixed_close_to_fp64(api, golden, case):
    S, po = api
    g, sysm, P, p = golden(case)
    print(f"igt tea mixed {case}", flush=True)
    a = S.BDSystem(sysm, P, precision=S.FP64, coord=_start(golden, case))
    b = S.BDSystem(sysm, P, precision=S.MIXED, coord=_start(golden, case))
    for k in range(3):
        b.set_coords(a.coords())
        b.set_rng_state(a.rng_state())
        c0 = a.coords()[:, :3].copy()
        a.compute_hydro(1)
        b.compute_hydro(1)
        da = a.coords()[:, :3] - c0
        db = b.coords()[:, :3] - c0
        assert np.linalg.norm(db - da) <= 1e-5 * np.linalg.norm(da), k


def test_igt_rejects_supplied_gaussians(api, golden):
    S, po = api
    g, sysm, P, p = golden("igt_const")
    s = S.BDSystem(sysm, P, precision=S.FP64)
    from bd_box_amd._lib import BackendError
    with pytest.raises(BackendError, match="device stream"):
        s.compute_hydro(1, np.zeros(3 * sysm.N))
