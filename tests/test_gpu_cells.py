"""Device cell lists: neighbour sets identical to brute force, forces equal to the
brute-force kernel's, trajectories equal to the oracle's.

  neighbour sets : bit-exact (same FP64 separation and predicate as the oracle's
                   brute-force pair finder)
  forces (FP64)  : <= 1e-12 relative to max|F| (summation order differs)
  forces (mixed) : <= 1e-5 relative to the FP64 brute-force forces
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


def _periodic_chain(N, box, seed):
    """A chain folded into a periodic cube (bonds cross the faces)."""
    sysm = wl.chain(N, seed=seed)
    c = sysm.coord.copy()
    c[:, :3] -= box * np.floor(c[:, :3] / box + 0.5)
    sysm.coord = c
    return sysm


@pytest.mark.parametrize("N,cutoffs", [(1, [5.0]), (2, [1.0, 100.0]), (400, [4.0, 12.01, 40.0])])
def test_neighbour_pairs_free_space(api, N, cutoffs):
    S, po = api
    sysm = wl.chain(max(N, 2), seed=11)
    sysm.coord = sysm.coord[:N].copy()
    sysm.Q, sysm.LJ = sysm.Q[:N].copy(), sysm.LJ[:N].copy()
    sysm.bonds = sysm.bonds[:0]
    sysm.bond_parms = sysm.bond_parms[:0]
    P = wl.parse_params(wl.chain_params(hydro="none"))
    s = S.BDSystem(sysm, P, precision=S.FP64)
    for cut in cutoffs:
        print(f"free space N={N} cutoff={cut}", flush=True)
        got = s.neighbour_pairs(cut)
        ref = po.pairs_within(sysm.coord, cut)
        assert np.array_equal(got, ref), (N, cut, len(got), len(ref))


def test_neighbour_pairs_flat_and_coincident(api):
    """A straight line (one cell across two axes) and beads sharing one position."""
    S, po = api
    N = 120
    coord = np.zeros((N, 4))
    coord[:, 0] = np.arange(N) * 3.8
    coord[:, 3] = 1.9
    coord[50:53, :3] = coord[50, :3]
    sysm = wl.System(coord, np.zeros(N), np.tile([2.0, 0.6], (N, 1)),
                     np.zeros((0, 2), np.int32), np.zeros((0, 3)))
    P = wl.parse_params(wl.chain_params(hydro="none", elec="no"))
    s = S.BDSystem(sysm, P, precision=S.FP64)
    for cut in (3.8, 7.7, 1000.0):
        got = s.neighbour_pairs(cut)
        assert np.array_equal(got, po.pairs_within(coord, cut)), cut


@pytest.mark.parametrize("cutoffs", [[7.0, 20.0, 29.0, 31.0, 80.0]])
def test_neighbour_pairs_periodic(api, cutoffs):
    """Cell counts per axis from 8 down to 1 (all distinct cells when fewer than 3)."""
    S, po = api
    box = 60.0
    sysm = _periodic_chain(300, box, seed=21)
    P = wl.parse_params(wl.chain_params(hydro="none", bc="periodic", xbox=box))
    s = S.BDSystem(sysm, P, precision=S.FP64)
    for cut in cutoffs:
        print(f"periodic box={box} cutoff={cut}", flush=True)
        got = s.neighbour_pairs(cut)
        ref = po.pairs_within(sysm.coord, cut, pbc=True, box=box)
        assert np.array_equal(got, ref), (cut, len(got), len(ref))


@pytest.mark.parametrize("geom", ["free", "periodic"])
def test_cell_forces_equal_brute_force(api, geom):
    S, po = api
    if geom == "free":
        sysm = wl.chain(600, seed=31)
        prm = dict(hydro="none", bond_lj_scale=0.5, bond_c_scale=0.25)
    else:
        sysm = _periodic_chain(400, 55.0, seed=32)
        prm = dict(hydro="none", bc="periodic", xbox=55.0)
    sysm.bond_parms[:, 1] = 40.0   # folded chains stretch a few bonds
    Pb = wl.parse_params(wl.chain_params(nb_list="brute", **prm))
    Pc = wl.parse_params(wl.chain_params(nb_list="bucket_space", **prm))
    Fo, Eo, _ = po.forces(sysm, po.params_from(Pb))
    F64b, E64b, rb = S.BDSystem(sysm, Pb, precision=S.FP64).compute_forces()
    F64c, E64c, rc = S.BDSystem(sysm, Pc, precision=S.FP64).compute_forces()
    assert _rel(F64c, F64b) < 1e-12
    assert _rel(F64c, Fo) < 1e-12
    np.testing.assert_allclose(E64c, E64b, rtol=1e-11, atol=1e-12)
    assert rb == rc
    F32c, E32c, _ = S.BDSystem(sysm, Pc, precision=S.MIXED).compute_forces()
    assert _rel(F32c, F64b) < 1e-5
    np.testing.assert_allclose(E32c, E64b, rtol=1e-5, atol=1e-8)


def test_cell_overlap_flag(api):
    """Two overlapping beads raise the repeat flag through the cell path as well."""
    S, _ = api
    sysm = wl.chain(40, seed=41)
    sysm.coord[20, :3] = sysm.coord[5, :3] + np.array([0.5, 0.0, 0.0])
    for nb in ("brute", "bucket_list"):
        P = wl.parse_params(wl.chain_params(hydro="none", check_overlap="yes", nb_list=nb))
        _, _, rep = S.BDSystem(sysm, P, precision=S.FP64).compute_forces()
        assert rep == 1, nb


@pytest.mark.parametrize("hydro", ["none", "cholesky"])
def test_cell_steps_vs_oracle(api, hydro):
    S, po = api
    N, steps = 90, 4
    sysm = wl.chain(N, seed=51)
    P = wl.parse_params(wl.chain_params(hydro=hydro, nb_list="bucket_space"))
    xi = np.random.default_rng(5).normal(size=(steps, 3 * N))
    ref = po.Sim(sysm, po.params_from(P), seed=1)
    ref.run(steps, xi)
    s = S.BDSystem(sysm, P, precision=S.FP64)
    s.compute_hydro(steps, xi)
    d_gpu = s.coords()[:, :3] - sysm.coord[:, :3]
    d_ref = ref.coords()[:, :3] - sysm.coord[:, :3]
    assert np.linalg.norm(d_gpu - d_ref) / np.linalg.norm(d_ref) < 1e-10


def test_neighbour_pairs_multi_kernel_build(api):
    """Above 65536 beads the cell list is built by the multi-kernel path (up to 48^3
    cells); the brute-force reference there is scipy's k-d tree, itself checked against the
    oracle's pair finder on the CPU (tests/test_pairs_oracle.py)."""
    from scipy.spatial import cKDTree
    S, _ = api
    N = 70000
    rng = np.random.default_rng(70000)
    coord = np.zeros((N, 4))
    coord[:, :3] = rng.uniform(-150.0, 150.0, size=(N, 3))
    coord[:, 3] = 1.0
    sysm = wl.System(coord, np.zeros(N), np.tile([0.5, 0.1], (N, 1)),
                     np.zeros((0, 2), np.int32), np.zeros((0, 3)))
    P = wl.parse_params(wl.chain_params(hydro="none", elec="no", cutoff_lj=2.5,
                                        nb_list="bucket_space"))
    s = S.BDSystem(sysm, P, precision=S.FP64)
    for cut in (2.5, 6.0):
        got = s.neighbour_pairs(cut)
        ref = np.array(sorted(cKDTree(coord[:, :3]).query_pairs(cut)), dtype=np.int64).reshape(-1, 2)
        assert np.array_equal(got, ref), (cut, len(got), len(ref))
