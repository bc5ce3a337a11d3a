"""A flow field on the device: the velocity gradient tensor G of the .prm keyword
vel_grad_tensor (diff_decl.c:72).  compute_hydro_ermak adds dt G r to every move and keeps
FS = dt G r + dt G r' (diff_equation.c:265-285); the IG-T corrector adds half of the
predictor's FS (:304-312).

  positions (FP64) : <= 1e-10 absolute against the reference's own trajectory
  dt               : identical (ermak_var rejections included)
  mixed precision  : per-step displacement within 1e-5 relative of the FP64 device run
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [("flow_ermak", 6), ("flow_geyer", 4), ("flow_igt", 6), ("flow_igt_retry", 8),
         ("flow_var_retry_none", 8)]


@pytest.fixture(scope="module")
def api():
    from bd_box_amd import system as S
    from oracle import pyoracle as po
    return S, po


@pytest.mark.parametrize("case,steps", CASES)
def test_flow_matches_reference(api, golden, case, steps):
    S, po = api
    g, sysm, P, p = golden(case)
    assert P.vel_grad_tensor is not None
    print(f"flow trajectory {case}", flush=True)
    s = S.BDSystem(sysm, P, precision=S.FP64)
    traj = g["traj.coord"].reshape(steps, -1, 4)
    for k in range(steps):
        s.compute_hydro(1)
        assert s.last_dt() == g["traj.dt"][k], (case, k)
        err = np.max(np.abs(s.coords()[:, :3] - traj[k][:, :3]))
        assert err < 1e-10, (case, k, err)
    sim = po.Sim(sysm, p, seed=P.rand_seed, algorithm=P.algorithm,
                 move_attempts=P.move_attempts, flow=P.vel_grad_tensor)
    sim.run(steps)
    assert s.rng_state() == sim.rng_state()


@pytest.mark.parametrize("case", ["flow_ermak", "flow_igt"])
def test_flow_off_is_plain_step(api, golden, case):
    """bdh_set_flow(NULL) returns to the plain move: the trajectory is the oracle's without
    a flow field."""
    S, po = api
    g, sysm, P, p = golden(case)
    print(f"flow off {case}", flush=True)
    s = S.BDSystem(sysm, P, precision=S.FP64)
    s.set_flow(None)
    sim = po.Sim(sysm, p, seed=P.rand_seed, algorithm=P.algorithm, move_attempts=P.move_attempts)
    for k in range(3):
        s.compute_hydro(1)
        sim.run(1)
        assert np.max(np.abs(s.coords()[:, :3] - sim.coords()[:, :3])) < 1e-10, k


@pytest.mark.parametrize("case,prec", [("flow_ermak", "MIXED"), ("flow_ermak", "MIXED_SPLIT"),
                                       ("flow_igt", "MIXED")])
def test_flow_mixed_close_to_fp64(api, golden, case, prec):
    S, po = api
    g, sysm, P, p = golden(case)
    print(f"flow mixed {case} {prec}", flush=True)
    a = S.BDSystem(sysm, P, precision=S.FP64)
    b = S.BDSystem(sysm, P, precision=getattr(S, prec))
    for k in range(4):
        b.set_coords(a.coords())
        b.set_rng_state(a.rng_state())
        c0 = a.coords()[:, :3].copy()
        a.compute_hydro(1)
        b.compute_hydro(1)
        da = a.coords()[:, :3] - c0
        db = b.coords()[:, :3] - c0
        assert np.linalg.norm(db - da) <= 1e-5 * np.linalg.norm(da), k


def test_flow_refused_by_reverse_step(api, golden):
    S, po = api
    from bd_box_amd._lib import BackendError
    g, sysm, P, p = golden("flow_ermak")
    s = S.BDSystem(sysm, P, precision=S.FP64)
    n = 3 * sysm.N
    with pytest.raises(BackendError, match="flow"):
        s.step_vjp(sysm.coord, np.zeros(n), np.ones(n))
