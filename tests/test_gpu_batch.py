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
