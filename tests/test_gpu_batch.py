"""Batches of independent chains (bdh_batch_*): every phase of the step is one launch over
all chains, each chain with its own launch shapes.

  per chain vs its own bdh_system (same precision, same seed or Gaussians): bit-identical
      positions and forces -- the batched kernels run the single-system kernel bodies
  FP64 batch vs the FP64 oracle (given Gaussians): <= 1e-10 relative displacement
  mixed batch vs FP64 batch: <= 1e-5 relative displacement
Chain sizes are uneven and cover one-tile matrices (N = 7: n = 21 padded to 128), several
outer panels (N = 130: n_pad = 512) and odd 3N (odd-length Gaussian tables).
"""
import numpy as np
import pytest

from bd_box_amd import workloads as wl

pytestmark = pytest.mark.gpu

SIZES = [20, 7, 130, 45, 64, 91]


@pytest.fixture(scope="module")
def api():
    from bd_box_amd import system as S
    from oracle import pyoracle as po
    return S, po


def _params(seed=1, **over):
    return wl.parse_params(wl.chain_params(rand_seed=seed, **over))


def _disp(c, c0):
    return c[:, :3] - c0[:, :3]


@pytest.mark.parametrize("prec", ["FP64", "MIXED", "MIXED_SPLIT"])
def test_batch_chains_match_own_systems_device_stream(api, prec):
    S, _ = api
    systems = wl.chains(SIZES, seed=500)
    seeds = [1000 + 17 * k for k in range(len(SIZES))]
    P = _params()
    refactor = 3
    print(f"batch vs own systems {prec}", flush=True)
    b = S.BDBatch(systems, P, seeds, precision=getattr(S, prec), refactor_every=refactor)
    b.compute_hydro(4)
    b.compute_hydro(3)
    F = b.forces()
    for c, (sysm, seed) in enumerate(zip(systems, seeds)):
        s = S.BDSystem(sysm, _params(seed), precision=getattr(S, prec), refactor_every=refactor)
        s.compute_hydro(4)
        s.compute_hydro(3)
        assert np.array_equal(b.chain_coords(c), s.coords()), (prec, c)
        o = 3 * b.offsets[c]
        assert np.array_equal(F[o:o + 3 * sysm.N], s.forces()), (prec, c)
        s.close()


def test_batch_edge_sizes(api):
    """Chains of 1, 2 and 3 beads (no bonds or one tile) beside a long one, and a batch of
    one chain: each equals its own system."""
    S, _ = api
    for sizes in ([1, 2, 3, 200], [57]):
        systems = wl.chains(sizes, seed=800)
        seeds = [31 + k for k in range(len(sizes))]
        print(f"batch edge sizes {sizes}", flush=True)
        b = S.BDBatch(systems, _params(), seeds, precision=S.MIXED_SPLIT, refactor_every=2)
        b.compute_hydro(5)
        for c, (sysm, seed) in enumerate(zip(systems, seeds)):
            s = S.BDSystem(sysm, _params(seed), precision=S.MIXED_SPLIT, refactor_every=2)
            s.compute_hydro(5)
            assert np.array_equal(b.chain_coords(c), s.coords()), (sizes, c)
            s.close()
        b.close()


def test_batch_given_noise_vs_oracle(api):
    S, po = api
    sizes = [12, 33, 8]
    systems = wl.chains(sizes, seed=600)
    P = _params()
    steps, refactor = 5, 2
    xi = np.random.default_rng(5).normal(size=(steps, 3 * sum(sizes)))
    print("batch given noise vs oracle", flush=True)
    b64 = S.BDBatch(systems, P, [1, 2, 3], precision=S.FP64, refactor_every=refactor)
    b64.compute_hydro(steps, xi)
    bmx = S.BDBatch(systems, P, [1, 2, 3], precision=S.MIXED_SPLIT, refactor_every=refactor)
    bmx.compute_hydro(steps, xi)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for c, sysm in enumerate(systems):
        xc = xi[:, 3 * off[c]:3 * off[c + 1]]
        ref = po.Sim(sysm, po.params_from(P), seed=1, refactor_every=refactor)
        ref.run(steps, xc)
        d_ref = _disp(ref.coords(), sysm.coord)
        for bb, tol in ((b64, 1e-10), (bmx, 1e-5)):
            d = _disp(bb.chain_coords(c), sysm.coord)
            err = np.linalg.norm(d - d_ref) / np.linalg.norm(d_ref)
            assert err < tol, (c, tol, err)


def test_batch_config_checks(api):
    S, _ = api
    from bd_box_amd._lib import BackendError
    systems = wl.chains([10, 12], seed=700)
    with pytest.raises(BackendError, match="hydro must be cholesky"):
        S.BDBatch(systems, _params(hydro="geyer"), [1, 2], precision=S.FP64)
    with pytest.raises(BackendError, match="check_overlap"):
        S.BDBatch(systems, _params(check_overlap="yes"), [1, 2], precision=S.FP64)


@pytest.mark.parametrize("left_max,outer", [(0, 256), (512, 256), (8192, 256), (-1, 512), (0, 1024),
                                            (8192, 1024)])
def test_batch_update_schedules(api, left_max, outer):
    """Every update schedule leaves the factor up to rounding: all right-looking (0), mixed
    (512: n_pad 128 and 512 chains left-looking, the rest right-looking) and all left-looking
    (8192, including a chain with n_pad 4224 that a single system factors right-looking),
    inside outer panels of 256, 512 or 1024 columns.
    FP64 vs the oracle (small chains) and vs the chain's own FP64 system (N = 1400):
    <= 1e-10 relative displacement; split-bf16 vs the oracle: <= 1e-5."""
    S, po = api
    sizes = [12, 130, 45, 1400]
    systems = wl.chains(sizes, seed=610)
    P = _params()
    steps, refactor = 3, 2
    xi = np.random.default_rng(6).normal(size=(steps, 3 * sum(sizes)))
    off = np.concatenate([[0], np.cumsum(sizes)])
    print(f"batch update schedule left_max={left_max} outer={outer}", flush=True)
    bb = {}
    for prec in ("FP64", "MIXED_SPLIT"):
        b = S.BDBatch(systems, P, [1, 2, 3, 4], precision=getattr(S, prec), refactor_every=refactor)
        b.set_tiles(left_max, outer)
        assert b.tiles() == (left_max, outer)
        b.compute_hydro(steps, xi)
        bb[prec] = b
    for c, sysm in enumerate(systems):
        xc = xi[:, 3 * off[c]:3 * off[c + 1]]
        if sysm.N <= 200:
            ref = po.Sim(sysm, po.params_from(P), seed=1, refactor_every=refactor)
            ref.run(steps, xc)
            tols = (("FP64", 1e-10), ("MIXED_SPLIT", 1e-5))
        else:
            ref = S.BDSystem(sysm, P, precision=S.FP64, refactor_every=refactor)
            ref.compute_hydro(steps, np.ascontiguousarray(xc))
            tols = (("FP64", 1e-10),)
        d_ref = _disp(ref.coords(), sysm.coord)
        for prec, tol in tols:
            d = _disp(bb[prec].chain_coords(c), sysm.coord)
            err = np.linalg.norm(d - d_ref) / np.linalg.norm(d_ref)
            assert err < tol, (c, prec, err)
    for b in bb.values():
        b.close()


def test_batch_autotune(api):
    """The autotuner times each candidate schedule, keeps the fastest, and the batch then
    steps exactly as one created with that schedule set by hand."""
    S, _ = api
    from bd_box_amd._lib import BackendError
    systems = wl.chains(SIZES, seed=500)
    P = _params()
    xi = np.random.default_rng(7).normal(size=(4, 3 * sum(SIZES)))
    print("batch autotune", flush=True)
    b = S.BDBatch(systems, P, [1] * len(SIZES), precision=S.MIXED_SPLIT, refactor_every=2)
    best, times = b.autotune(reps=2)
    n_pads = sorted({-(-3 * n // 128) * 128 for n in SIZES})
    assert 0 in times and n_pads[-1] in times and best in times
    assert set(times) <= {0, *n_pads}
    assert all(t > 0 for t in times.values())
    assert times[best] == min(times.values())
    assert b.schedule() == best
    b.compute_hydro(4, xi)
    r = S.BDBatch(systems, P, [1] * len(SIZES), precision=S.MIXED_SPLIT, refactor_every=2)
    r.set_schedule(best)
    r.compute_hydro(4, xi)
    assert np.array_equal(b.coords(), r.coords())
    with pytest.raises(BackendError, match="before the first step"):
        b.autotune()
    b.close()
    r.close()


@pytest.mark.parametrize("prec", ["FP64", "MIXED"])
def test_batch_cycle_graph_replays_match_launches(api, prec):
    """A whole refactor cycle of device-stream steps is one captured graph: three replays,
    with the host reading the state between them, give bitwise the positions and forces of
    the same steps run as direct launches (phase timing on), cycle after cycle; a partial
    cycle after them takes the launch path."""
    S, _ = api
    systems = wl.chains(SIZES, seed=510)
    seeds = [2000 + 13 * k for k in range(len(SIZES))]
    P = _params()
    R = 4
    print(f"batch cycle graph {prec}", flush=True)
    g = S.BDBatch(systems, P, seeds, precision=getattr(S, prec), refactor_every=R)
    d = S.BDBatch(systems, P, seeds, precision=getattr(S, prec), refactor_every=R)
    d.phase_timing(True)
    for cycle in range(3):
        g.compute_hydro(R)
        d.compute_hydro(R)
        assert np.array_equal(g.coords(), d.coords()), (prec, cycle)
        assert np.array_equal(g.forces(), d.forces()), (prec, cycle)
    g.compute_hydro(2)
    d.compute_hydro(2)
    assert np.array_equal(g.coords(), d.coords())
    g.close()
    d.close()


def test_batch_autotune_tiles(api):
    """The tile tuner times every (schedule, outer width) candidate, keeps the fastest, and
    the batch then steps exactly as one with those tiles set by hand."""
    S, _ = api
    systems = wl.chains([300, 700, 90], seed=520)   # n_pad 1024 and 2176 reach widths 512, 1024
    P = _params()
    xi = np.random.default_rng(8).normal(size=(4, 3 * 1090))
    print("batch autotune tiles", flush=True)
    b = S.BDBatch(systems, P, [1, 2, 3], precision=S.MIXED, refactor_every=2)
    best, times = b.autotune_tiles(reps=2)
    assert {o for _, o in times} == {256, 512, 1024}
    assert best in times and times[best] == min(times.values())
    assert b.tiles() == best
    b.compute_hydro(4, xi)
    r = S.BDBatch(systems, P, [1, 2, 3], precision=S.MIXED, refactor_every=2)
    r.set_tiles(*best)
    r.compute_hydro(4, xi)
    assert np.array_equal(b.coords(), r.coords())
    b.close()
    r.close()
