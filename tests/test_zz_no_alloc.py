"""The allocation-free Gaussian entry point against the allocating one, through the C ABI.

bdh_gauss_noalloc stages through a fixed device array instead of hipMalloc'd buffers; for the
same stream state and count its Gaussians and the updated 48-bit state must be bit-identical
to bdh_gauss's (src/rand_move.c:132-168 is the stream both reproduce).  Two calls in a row,
the second smaller, check that nothing of the first call's contents leaks into the second.
"""
import ctypes as C

import numpy as np
import pytest

from bd_box_amd import _lib

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 20   # BDH_GAUSS_NOALLOC_MAX in include/bdbox_hip.h


def _draw(fn, state, count):
    L = _lib.load()
    st = C.c_uint64(state)
    out = np.full(count, np.nan)
    _lib.check(getattr(L, fn)(C.byref(st), out.ctypes.data_as(_lib._dp), count), fn)
    return out, st.value


@pytest.fixture(scope="module")
def lib():
    L = _lib.load()
    _lib.check(L.bdh_init(0), "bdh_init")
    return L


@pytest.mark.parametrize("seed, counts", [(12345, (30001, 777)), (7, (NOALLOC_MAX, 3))])
def test_noalloc_gauss_matches_gauss(lib, seed, counts):
    s_ref = s_new = lib.bdh_seed_state(seed)
    for count in counts:
        ref, s_ref = _draw("bdh_gauss", s_ref, count)
        new, s_new = _draw("bdh_gauss_noalloc", s_new, count)
        assert s_new == s_ref
        assert np.array_equal(new, ref)          # bit-identical, including the odd tail
        assert np.all(np.isfinite(new))


def test_noalloc_gauss_refuses_oversize(lib):
    st = C.c_uint64(lib.bdh_seed_state(1))
    out = np.zeros(1)
    rc = lib.bdh_gauss_noalloc(C.byref(st), out.ctypes.data_as(_lib._dp), NOALLOC_MAX + 1)
    assert rc < 0
    assert b"BDH_GAUSS_NOALLOC_MAX" in lib.bdh_last_error()
    assert st.value == lib.bdh_seed_state(1)    # state untouched on refusal
