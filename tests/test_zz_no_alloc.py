"""GPU: bsn_conserved_totals_noalloc (the allocation-free variant of
bsn_conserved_totals, whose partial sums sit in a fixed device array) against
bsn_conserved_totals through the C ABI: the same sums bit for bit on two
calls in a row, the second on a smaller grid, with the sums returned to host
and to device memory."""
import ctypes as C

import numpy as np
import pytest

import binary_supernova_amd as bsn
from binary_supernova_amd._lib import BSN_DEVICE, BSN_HOST, call

pytestmark = pytest.mark.gpu


def state(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    U = rng.uniform(0.1, 2.0, (nz + 6, ny + 6, nx + 6, 5))
    U[..., 1:4] -= 1.0
    return U


def totals(g, dV):
    out = (C.c_double * 5)()
    call("bsn_conserved_totals", g.handle, float(dV), out)
    return np.array(list(out))


def totals_noalloc_host(g, dV):
    out = np.full(5, np.nan)
    call("bsn_conserved_totals_noalloc", g.handle, float(dV), C.c_void_p(out.ctypes.data), BSN_HOST)
    return out


def hip():
    # the HIP runtime libbsn_hip.so itself links (the test allocates the
    # device output buffer; the entry point under test does not)
    L = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return L


def totals_noalloc_device(g, dV):
    L = hip()
    d = C.c_void_p()
    assert L.hipMalloc(C.byref(d), 5 * 8) == 0
    try:
        fill = np.full(5, np.nan)
        assert L.hipMemcpy(d, C.c_void_p(fill.ctypes.data), 40, 1) == 0     # host to device
        call("bsn_conserved_totals_noalloc", g.handle, float(dV), d, BSN_DEVICE)
        g.ctx.synchronize()
        out = np.empty(5)
        assert L.hipMemcpy(C.c_void_p(out.ctypes.data), d, 40, 2) == 0      # device to host
    finally:
        L.hipFree(d)
    return out


def test_noalloc_totals_equal_the_allocating_entry_point():
    print("conserved totals without allocation", flush=True)
    for (nx, ny, nz), seed, dV in (((96, 80, 72), 1, 0.125), ((20, 9, 13), 2, 0.5)):
        g = bsn.Grid(nx, ny, nz).set_state(state(nx, ny, nz, seed))
        want = totals(g, dV)
        np.testing.assert_array_equal(totals_noalloc_host(g, dV), want)
        np.testing.assert_array_equal(totals_noalloc_device(g, dV), want)
        # and they are the sums of the active cells
        U = g.get_state()[3:-3, 3:-3, 3:-3]
        np.testing.assert_allclose(want, U.sum((0, 1, 2)) * dV, rtol=1e-12)


def test_noalloc_totals_reject_bad_arguments():
    g = bsn.Grid(8, 8, 8).set_state(state(8, 8, 8, 3))
    out = np.zeros(5)
    with pytest.raises(bsn.BsnError, match="where"):
        call("bsn_conserved_totals_noalloc", g.handle, 1.0, C.c_void_p(out.ctypes.data), 7)
    with pytest.raises(bsn.BsnError, match="NULL"):
        call("bsn_conserved_totals_noalloc", g.handle, 1.0, None, BSN_HOST)
