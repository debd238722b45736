"""The forward that allocates nothing (prc_projector_forward_noalloc), through the C ABI, against
the allocating device-pointer forward (prc_projector_forward): bit-identical results on two calls
in a row, the second on a smaller volume, and a clear error for a volume beyond its scratch."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import announce
from prc_mlem_amd import phantoms

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


def hip():
    for name in ("libamdhip64.so", os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")):
        try:
            L = C.CDLL(name)
            break
        except OSError:
            continue
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


class Device:
    """A device buffer of float32 holding one volume."""

    def __init__(self, H, shape):
        self.H, self.shape, self.nbytes = H, shape, 4 * int(np.prod(shape))
        self.ptr = C.c_void_p()
        assert H.hipMalloc(C.byref(self.ptr), self.nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert self.H.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape, np.float32)
        assert self.H.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.ptr)


def both_forwards(prc, A, x):
    L, H = prc.lib(), hip()
    noalloc = L.prc_projector_forward_noalloc
    noalloc.restype, noalloc.argtypes = C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]
    din, d1, d2 = Device(H, x.shape), Device(H, x.shape), Device(H, x.shape)
    try:
        din.put(x)
        d1.put(np.full(x.shape, np.nan, np.float32))  # every voxel must be written
        d2.put(np.full(x.shape, np.nan, np.float32))
        assert noalloc(A.handle, din.ptr, d1.ptr) == 0, L.prc_last_error().decode()
        assert L.prc_projector_forward(A.handle, din.ptr, d2.ptr) == 0, L.prc_last_error().decode()
        assert H.hipDeviceSynchronize() == 0
        return d1.get(), d2.get()
    finally:
        din.free(), d1.free(), d2.free()


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_noalloc_forward_equals_forward_twice(prc, precision):
    rho, truth = phantoms.cylinder_phantom()
    announce(f"no-allocation forward, 64^3 cylinder 68Ga, {precision}")
    A = prc.Projector(rho, (2.0, 2.0, 2.0), "68Ga", 64, precision=precision)
    a, b = both_forwards(prc, A, truth)
    np.testing.assert_array_equal(a, b)
    assert np.isfinite(a).all()
    np.testing.assert_allclose(a, A.forward(truth), rtol=0, atol=0)
    A.close()
    announce(f"no-allocation forward, 16x14x12 test phantom, {precision} (smaller, second call)")
    rho2, x2, _, sp2 = phantoms.test_phantom()
    B = prc.Projector(rho2, sp2, "68Ga", 32, precision=precision)
    a2, b2 = both_forwards(prc, B, x2)
    np.testing.assert_array_equal(a2, b2)
    np.testing.assert_array_equal(a2, B.forward(x2))
    B.close()


def test_noalloc_forward_rejects_a_volume_beyond_its_scratch(prc):
    # 400 x 400 x 210 voxels = 33.6 M floats before margins, above the 32 Mi-float scratch.
    shape = (210, 400, 400)
    rho = np.full(shape, 0.0012, np.float32)
    rho[100:104, 200:204, 200:204] = 1.0
    announce("no-allocation forward: oversize volume is rejected")
    A = prc.Projector(rho, (2.0, 2.0, 2.0), "18F", 8)
    L = prc.lib()
    noalloc = L.prc_projector_forward_noalloc
    noalloc.restype, noalloc.argtypes = C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]
    H = hip()
    d = Device(H, shape)
    try:
        rc = noalloc(A.handle, d.ptr, d.ptr)
        assert rc == -1
        assert "no-allocation forward holds" in L.prc_last_error().decode()
    finally:
        d.free()
        A.close()
