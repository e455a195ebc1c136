"""mhd_expand_noalloc (device arrays, fixed scratch, no allocation) against
mhd_expand (host arrays) through the C ABI: two calls in a row, the second
on a smaller box, and the size limit.  Device memory for the test's own
inputs and outputs comes from the HIP runtime the library itself links."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = ctypes.CDLL("libamdhip64.so.7")   # the runtime libmhd_hip.so is linked against
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    for f in (hip.hipMalloc, hip.hipFree, hip.hipMemcpy, hip.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return hip


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape):
        out = np.empty(shape)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = ctypes.c_void_p()


def test_expand_noalloc_matches_expand_twice():
    import fug30v0kxc_amd as fa
    from fug30v0kxc_amd import _abi
    L = _abi.load()
    hip = _hip()
    rng = np.random.RandomState(5)
    boxes = [(64, 48, 20), (32, 16, 6)]   # the second call on a smaller box
    bufs = []
    try:
        cap_in = max(nx * ny * nz for nx, ny, nz in boxes) * 8
        din, dout = DevBuf(hip, cap_in), DevBuf(hip, 4 * cap_in)
        bufs += [din, dout]
        for nx, ny, nz in boxes:
            print(f"[no-alloc] expand {nx}x{ny}x{nz}", flush=True)
            f = rng.randn(nz, ny, nx)
            ref = fa.expand(f)
            din.put(f)
            rc = L.mhd_expand_noalloc(nx, ny, nz, din.p, dout.p, None)
            assert rc == 0, rc
            assert hip.hipDeviceSynchronize() == 0
            got = dout.get((nz, 2 * ny, 2 * nx))
            assert np.array_equal(got, ref)
            # the input was only read
            assert np.array_equal(din.get((nz, ny, nx)), f)
    finally:
        for b in bufs:
            b.free()


def test_expand_noalloc_rejects_boxes_beyond_the_scratch():
    from fug30v0kxc_amd import _abi
    L = _abi.load()
    # 2 nx ny nz + 6 (nx + ny) doubles must fit in the scratch: 256^2 x 64 does not
    nx, ny, nz = 256, 256, 64
    assert 8 * (2 * nx * ny * nz + 6 * (nx + ny)) > _abi.MHD_NOALLOC_SCRATCH_BYTES
    # the size check comes before any device access; the pointers are never used
    rc = L.mhd_expand_noalloc(nx, ny, nz, ctypes.c_void_p(256), ctypes.c_void_p(256), None)
    assert rc == _abi.MHD_E_SCRATCH
    assert L.mhd_strerror(rc) == b"box too large for the no-allocation scratch"
    assert L.mhd_expand_noalloc(nx, ny, nz, None, None, None) == _abi.MHD_E_ARG
