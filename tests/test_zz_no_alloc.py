"""hbem_rsvd_sv_noalloc: randomized_svd::sv on device-resident inputs with
every work buffer in the library's fixed per-GPU array (no device
allocation).  Its sigma must match hbem_rsvd_sv's on the same matrix and test
matrix to the sigma bar (1e-10 relative), on two calls in a row where the
second is smaller (the array keeps the first call's contents), and a size
beyond the array's limit must be refused before anything is enqueued.

Device memory for the inputs is allocated here through the HIP runtime the
library itself loaded (ctypes), so pointers and kernels share one runtime."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip_runtime():
    from helmholtzbem_amd import _lib
    _lib.lib()
    rt = ctypes.CDLL("libamdhip64.so.7")      # the already loaded runtime (matched by soname)
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


class DevArray:
    def __init__(self, rt, nbytes):
        self.rt = rt
        self.p = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a) if a.flags.c_contiguous else np.asfortranarray(a)
        assert a.nbytes == self.nbytes
        assert self.rt.hipMemcpy(self.p, ctypes.c_void_p(a.ctypes.data), a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        out = np.zeros(shape, dtype=dtype)
        assert self.rt.hipMemcpy(ctypes.c_void_p(out.ctypes.data), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p.value:
            self.rt.hipFree(self.p)
            self.p = ctypes.c_void_p()


def noalloc_sv(H, rt, T, R, q):
    from helmholtzbem_amd import _lib
    n, nc = T.shape[0], R.shape[1]
    dT, dR, dS = DevArray(rt, T.nbytes), DevArray(rt, R.nbytes), DevArray(rt, 8)
    try:
        dT.put(np.asfortranarray(T))
        dR.put(np.asfortranarray(R))
        print("noalloc call n=%d nc=%d q=%d" % (n, nc, q), flush=True)
        _lib.check(_lib.lib().hbem_rsvd_sv_noalloc(dT.p, n, dR.p, nc, q, dS.p, None))
        assert rt.hipDeviceSynchronize() == 0
        return float(dS.get(1, np.float64)[0])
    finally:
        for d in (dT, dR, dS):
            d.free()


def test_noalloc_matches_rsvd_sv_twice_then_smaller(H):
    rt = hip_runtime()
    rng = np.random.default_rng(23)
    for n, nc, q in ((1500, 2, 2), (400, 3, 1)):
        T = np.asfortranarray(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        R = H.gaussian_test_matrix(n, nc, n)
        ref = H.randomized_svd.sv(T, R, q)
        s = noalloc_sv(H, rt, T, R, q)
        assert abs(s - ref) <= 1e-10 * ref, (n, s, ref)


def test_noalloc_on_a_transmission_matrix(H):
    rt = hip_runtime()
    mesh = H.square_mesh(0.5, 12)
    sp = H.ContinuousSpace1()
    bd = H.BuilderData(mesh, sp, sp, 11)
    T = H.SolutionsOperator(bd).gen_sol_op(2.0, 1.0, 20.0)
    R = H.gaussian_test_matrix(96, 2, 0)
    assert abs(noalloc_sv(H, rt, T, R, 2) - H.randomized_svd.sv(T, R, 2)) <= 1e-10 * H.randomized_svd.sv(T, R, 2)


def test_noalloc_rejects_sizes_beyond_the_array(H):
    from helmholtzbem_amd import _lib
    rt = hip_runtime()
    d = DevArray(rt, 256)
    try:
        rc = _lib.lib().hbem_rsvd_sv_noalloc(d.p, 2049, d.p, 2, 1, d.p, None)
        assert rc == _lib.HBEM_EINVAL
        buf = ctypes.create_string_buffer(256)
        _lib.lib().hbem_last_error(buf, 256)
        assert b"no-allocation limit" in buf.value
        rc = _lib.lib().hbem_rsvd_sv_noalloc(d.p, 16, d.p, 9, 1, d.p, None)
        assert rc == _lib.HBEM_EINVAL
    finally:
        d.free()
