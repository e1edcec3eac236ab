"""gzm_riemann_states_noalloc (face states staged through the library's fixed
device array, no device allocation) against gzm_riemann_states (buffers
allocated per call) through the C ABI: the same kernel, so the same bits.
Two calls in a row, the second smaller, so a result left in the shared
array by the first call cannot pass for the second's; then the size limit."""
import ctypes as C

import numpy as np
import pytest

import gzm

pytestmark = pytest.mark.gpu

NMAX = 65536  # GZM_RIEMANN_NOALLOC_MAX


def states(n, seed):
    rng = np.random.default_rng(seed)
    L = np.column_stack([rng.uniform(0.1, 3, n), rng.uniform(0.05, 5, n), rng.normal(size=(n, 3))])
    R = np.column_stack([rng.uniform(0.1, 3, n), rng.uniform(0.05, 5, n), rng.normal(size=(n, 3))])
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    # limiters spread over the cascade: HLLC stands, Roe / PVRS estimates, exact solver
    lim = rng.uniform(0.05, 8.0, n)
    return L, R, nrm, lim


@pytest.mark.parametrize("mode,precision", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_noalloc_matches_allocating_entry_on_two_calls(mode, precision):
    print(f"no-alloc Riemann states: mode {mode} precision {precision}", flush=True)
    for n, seed in ((NMAX, 11 + mode + 2 * precision), (3001, 17 + mode + 2 * precision)):
        L, R, nrm, lim = states(n, seed)
        lim_arg = lim if mode == 0 else None
        ref, ref_mask = gzm.riemann_states(L, R, nrm, lim_arg, mode=mode, precision=precision)
        out, mask = gzm.riemann_states(L, R, nrm, lim_arg, mode=mode, precision=precision, noalloc=True)
        np.testing.assert_array_equal(mask, ref_mask, err_msg=f"n={n}")
        np.testing.assert_array_equal(out, ref, err_msg=f"n={n}")
        assert np.all(np.isfinite(out))


def test_noalloc_rejects_more_than_its_array_holds():
    L, R, nrm, lim = states(NMAX + 1, 5)
    with pytest.raises(RuntimeError, match="out of range"):
        gzm.riemann_states(L, R, nrm, lim, noalloc=True)
    lib = gzm.load_library()
    # the limit is checked before any pointer is read or any device work
    dummy = np.zeros(1)
    p = dummy.ctypes.data_as(C.POINTER(C.c_double))
    m = np.zeros(1, dtype=np.int32).ctypes.data_as(C.POINTER(C.c_int32))
    assert lib.gzm_riemann_states_noalloc(0, 0, 0, 5.0 / 3.0, NMAX + 1, p, p, p, p, p, m) != 0
