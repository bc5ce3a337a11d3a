"""Edge sizes through the C ABI against the FP64 oracle: one bead (3 x 3 tensor, padded to
one 64-column panel; the TEA sums have no pairs, so eps = 0 and beta = 1/2 as in
diff_rand.c:207-212), two beads, and 3N on either side of the 64-bead tiles and 64-column
panels (N = 21, 22: n = 63, 66; N = 64, 65, 129).

  FP64  : positions <= 1e-10 relative displacement, forces <= 1e-9 relative
  mixed : displacements <= 1e-5 relative
"""
import numpy as np
import pytest

from bd_box_amd import workloads as wl

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 21, 22, 64, 65, 129]


@pytest.fixture(scope="module")
def api():
    from bd_box_amd import system as S
    from oracle import pyoracle as po
    return S, po


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("hydro", ["cholesky", "geyer", "none"])
@pytest.mark.parametrize("N", SIZES)
def test_edge_sizes_steps_vs_oracle(api, hydro, N):
    S, po = api
    sysm = wl.chain(N, seed=500 + N)
    P = wl.parse_params(wl.chain_params(hydro=hydro))
    xi = np.random.default_rng(N).normal(size=(3, 3 * N))
    print(f"edge size N={N} hydro={hydro}", flush=True)
    ref = po.Sim(sysm, po.params_from(P), seed=1)
    ref.run(3, xi)
    d_ref = ref.coords()[:, :3] - sysm.coord[:, :3]
    for prec, tol in ((S.FP64, 1e-10), (S.MIXED, 1e-5)):
        s = S.BDSystem(sysm, P, precision=prec)
        s.compute_hydro(3, xi)
        d = s.coords()[:, :3] - sysm.coord[:, :3]
        assert _rel(d, d_ref) < tol, (prec, _rel(d, d_ref))
        if prec == S.FP64 and N > 1:
            assert _rel(s.forces(), ref.forces()) < 1e-9
        s.close()


@pytest.mark.parametrize("N", [1, 2, 65])
def test_edge_sizes_igt_vs_oracle(api, N):
    """IG-T (stream noise) at the same edge sizes, Cholesky and diagonal noise."""
    S, po = api
    for hydro in ("cholesky", "none"):
        sysm = wl.chain(N, seed=600 + N)
        P = wl.parse_params(wl.chain_params(hydro=hydro, algorithm="igt_const", rand_seed=17))
        print(f"edge size igt N={N} hydro={hydro}", flush=True)
        ref = po.Sim(sysm, po.params_from(P), seed=P.rand_seed, algorithm=P.algorithm)
        ref.run(3)
        s = S.BDSystem(sysm, P, precision=S.FP64)
        s.compute_hydro(3)
        assert s.rng_state() == ref.rng_state()
        assert np.max(np.abs(s.coords()[:, :3] - ref.coords()[:, :3])) < 1e-10
        s.close()
