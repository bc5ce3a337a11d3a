"""Rejected steps on the device: comp_step_ermak's loop (diff_comp.c:152-179).

With check_overlap a step whose new positions overlap is undone and redrawn with the next
Gaussians of the stream; ermak_var halves dt for the first 16 rejections.  The device path
must take the same decisions as the reference, so the trajectories and the per-step dt are
compared with the reference's own (tests/golden/retry_*.npz, written by the compiled
reference) and with the oracle.

  positions (FP64) : <= 1e-10 absolute after 8 steps (the reference's are ~1e1)
  dt               : identical
"""
import numpy as np
import pytest

from bd_box_amd import workloads as wl

pytestmark = pytest.mark.gpu

CASES = ["retry_const", "retry_var", "retry_var_none", "retry_var_geyer"]


@pytest.fixture(scope="module")
def api():
    from bd_box_amd import system as S
    from oracle import pyoracle as po
    return S, po


@pytest.mark.parametrize("case", CASES)
def test_rejected_steps_match_reference(api, golden, case):
    S, po = api
    g, sysm, P, p = golden(case)
    print(f"retry trajectory {case}", flush=True)
    s = S.BDSystem(sysm, P, precision=S.FP64)
    traj = g["traj.coord"].reshape(8, -1, 4)
    for k in range(8):
        s.compute_hydro(1)
        assert s.last_dt() == g["traj.dt"][k], (case, k, s.last_dt(), g["traj.dt"][k])
        err = np.max(np.abs(s.coords()[:, :3] - traj[k][:, :3]))
        assert err < 1e-10, (case, k, err)
    # the device stream advanced by every redraw, as the reference's did
    sim = po.Sim(sysm, p, seed=P.rand_seed, algorithm=P.algorithm,
                 move_attempts=P.move_attempts)
    sim.run(8)
    assert s.rng_state() == sim.rng_state()
    assert sim.rejections() > 8


def test_supplied_gaussians_then_stream_redraws(api, golden):
    """xi supplies each step's first draw; redraws come from the device stream (the
    oracle's rule for the same call)."""
    S, po = api
    g, sysm, P, p = golden("retry_var")
    xi = np.random.default_rng(3).normal(size=(4, 3 * sysm.N))
    print("retry with supplied xi", flush=True)
    s = S.BDSystem(sysm, P, precision=S.FP64)
    s.compute_hydro(4, xi)
    sim = po.Sim(sysm, p, seed=P.rand_seed, algorithm=P.algorithm,
                 move_attempts=P.move_attempts)
    sim.run(4, xi)
    assert sim.rejections() > 0
    assert s.last_dt() == sim.last_dt()
    assert np.max(np.abs(s.coords()[:, :3] - sim.coords()[:, :3])) < 1e-10
    assert s.rng_state() == sim.rng_state()


def test_attempt_limit_is_fatal(api, golden):
    S, po = api
    g, sysm, P, p = golden("retry_const")
    prm = {k[4:]: str(v) for k, v in g.items() if k.startswith("prm.")}
    prm["move_attempts"] = "0"
    P0 = wl.parse_params(prm)
    print("retry attempt limit", flush=True)
    s = S.BDSystem(sysm, P0, precision=S.FP64)
    from bd_box_amd._lib import BackendError
    with pytest.raises(BackendError, match="Wrong conformation"):
        s.compute_hydro(8)


@pytest.mark.parametrize("prec", ["MIXED", "MIXED_SPLIT"])
def test_mixed_precision_retries_run(api, golden, prec):
    """Mixed precision takes its own decisions near the overlap threshold; it must still
    redraw rather than accept an overlapping step: no accepted configuration overlaps."""
    S, po = api
    g, sysm, P, p = golden("retry_var")
    print(f"retry mixed {prec}", flush=True)
    s = S.BDSystem(sysm, P, precision=getattr(S, prec))
    for k in range(8):
        s.compute_hydro(1)
        c = s.coords()
        _, _, rep = po.forces(sysm, p, coord=c)
        assert rep == 0, k
        assert s.last_dt() in (2.0, 1.0, 0.5, 0.25, 0.125, 0.0625, 0.03125, 0.015625)
