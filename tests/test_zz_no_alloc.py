"""amt_oc_update_noalloc (the OC update on caller device buffers, scratch in a static device
array) against amt_oc_update, the handle's entry point with its own buffers: same bisection
pass count, bitwise-equal densities and change, on two calls in a row, the second with a
smaller input. The caller's device buffers come from hipMalloc of the HIP runtime libamt.so
links; the entry point itself allocates nothing."""
import ctypes as C

import numpy as np
import pytest

import am_ml_suite_amd as A
from am_ml_suite_amd import _ffi

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class _Dev:
    """One device buffer holding a copy of a host array."""

    def __init__(self, hip, arr):
        self.hip, self.n = hip, arr.nbytes
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(self.n, 8)) == 0
        assert hip.hipMemcpy(self.ptr, arr.ctypes.data, self.n, H2D) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.n, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _case(nx, ny, nz, vf, seed):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    rho = rng.uniform(0.05, 0.95, n)
    rho[: n // 20] = 1.0          # non-design (>= 0.99)
    rho[n // 20: n // 12] = 0.01  # non-design (void)
    dcf = -rng.uniform(0.001, 4.0, n)
    dcf[n // 3] = 0.5             # positive: the NaN branch
    return rho, dcf


def _with_handle(nx, ny, nz, rho, dcf, vf, move, rho_min):
    c = A.SIMPConfig(verbose=False, volume_fraction=vf, move_limit=move, rho_min=rho_min)
    opt = A.TopologyOptimizer(nx, ny, nz, 1.0, A.MaterialProperties(1.0, 0.3), c)
    opt.set_density(rho)
    opt.begin_run()
    ch, passes = opt.oc_update(dcf)
    return opt.get_density(), ch, passes


def _noalloc(L, hip, rho, dcf, design, vf, move, rho_min):
    d_rho, d_dcf, d_des = _Dev(hip, rho), _Dev(hip, dcf), _Dev(hip, design)
    d_res = _Dev(hip, np.zeros(4))
    try:
        rc = L.amt_oc_update_noalloc(d_rho.ptr, d_dcf.ptr, d_des.ptr, len(rho), vf, move,
                                     rho_min, d_res.ptr, None)
        assert rc == 0, L.amt_last_error(None)
        assert hip.hipDeviceSynchronize() == 0
        return d_rho.get(rho), d_res.get(np.zeros(4))
    finally:
        for d in (d_rho, d_dcf, d_des, d_res):
            d.free()


def test_oc_update_noalloc_matches_handle_entry_point():
    L, hip = _ffi.load(), _hip()
    move, rho_min = 0.15, 0.01
    # the second call is smaller: nothing of the first call's scratch may leak into it
    for (nx, ny, nz, vf, seed) in [(40, 30, 20, 0.45, 8), (12, 9, 7, 0.35, 9)]:
        rho, dcf = _case(nx, ny, nz, vf, seed)
        design = ((rho > rho_min + 1e-6) & (rho < 0.99)).astype(np.uint8)
        ref, ch, passes = _with_handle(nx, ny, nz, rho, dcf, vf, move, rho_min)
        got, res = _noalloc(L, hip, rho, dcf, design, vf, move, rho_min)
        print("%dx%dx%d vf %g: passes %d / %d, change %.17g / %.17g"
              % (nx, ny, nz, vf, res[2], passes, res[0], ch), flush=True)
        assert int(res[2]) == passes
        assert int(res[3]) == int(design.sum())
        assert np.array_equal(got, ref)                 # same kernels, same order: bitwise
        assert res[0] == ch
        assert res[1] == pytest.approx(got[design == 1].sum(), rel=1e-13)


def test_oc_update_noalloc_rejects_bad_sizes():
    L = _ffi.load()
    x = C.c_void_p(16)  # never dereferenced: the size check comes first
    assert L.amt_oc_update_noalloc(x, x, x, 0, 0.3, 0.2, 0.01, x, None) == _ffi.AMT_EINVAL
    assert L.amt_oc_update_noalloc(x, x, x, (1 << 31) + 1, 0.3, 0.2, 0.01, x,
                                   None) == _ffi.AMT_EINVAL
    assert b"2^31" in L.amt_last_error(None)
    assert L.amt_oc_update_noalloc(None, x, x, 10, 0.3, 0.2, 0.01, x, None) == _ffi.AMT_EINVAL
