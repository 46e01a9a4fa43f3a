"""The allocation-free mass / momentum reduction against the allocating one,
through the C ABI: the same bits, on two calls in a row, the second on a
smaller block (the shared scratch keeps the first call's partials, which the
second must not see)."""
import ctypes as C

import numpy as np
import pytest

from freelb_amd import abi
from freelb_amd.cases import cavity_block

pytestmark = pytest.mark.gpu

HIP_D2H = 2  # hipMemcpyDeviceToHost


@pytest.mark.parametrize("storage", ["f32", "f16s", "f64"])
def test_noalloc_reduction_equals_allocating_one(storage):
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    d_out = C.c_void_p()
    assert hip.hipMalloc(C.byref(d_out), 4 * 8) == 0
    L = abi.lib()
    try:
        for shape, steps, mask in (((40, 33, 21), 7, 0xFE), ((9, 6, 5), 4, 0x02)):
            print(f"[noalloc] {storage} {shape} steps {steps} mask {mask:#x}", flush=True)
            b = cavity_block(shape, storage=storage, u0_amp=1e-3, seed=3)
            try:
                b.step(steps)
                want = b.mass_momentum(mask)
                abi.check(L.flb_reduce_mass_momentum_noalloc(b.handle, mask, d_out))
                b.synchronize()
                got = np.zeros(4)
                assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), d_out, 4 * 8, HIP_D2H) == 0
                assert np.array_equal(got, want), (got, want)
                assert got[0] > 0
            finally:
                b.close()
    finally:
        hip.hipFree(d_out)
