"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this, and only as the checker / CPU baseline. The product never imports it.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")


class OrcConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("h", C.c_double), ("E", C.c_double), ("nu", C.c_double),
        ("volume_fraction", C.c_double), ("penalty", C.c_double),
        ("filter_radius", C.c_double), ("move_limit", C.c_double),
        ("rho_min", C.c_double), ("E_min", C.c_double),
        ("max_iterations", C.c_int), ("convergence_tol", C.c_double),
        ("pcg_tol", C.c_double), ("pcg_maxit", C.c_int),
        ("penalty_factor", C.c_double), ("ref_compat", C.c_int), ("verbose", C.c_int),
    ]


class OrcIterResult(C.Structure):
    _fields_ = [
        ("compliance", C.c_double), ("volume_fraction", C.c_double),
        ("max_density_change", C.c_double),
        ("solver_iterations", C.c_int), ("solver_converged", C.c_int),
        ("bisect_passes", C.c_int),
        ("true_rel_residual", C.c_double),
        ("assembly_time_s", C.c_double), ("solve_time_s", C.c_double),
        ("sensitivity_time_s", C.c_double), ("update_time_s", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.orc_ke0.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.orc_element_dofs.argtypes = [C.c_int] * 6 + [_i64p]
        L.orc_assemble_csr.restype = C.c_int64
        L.orc_assemble_csr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_double, _dp, C.c_double, C.c_double, C.c_double,
                                       C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_csr_spmv.argtypes = [C.c_int64, _i64p, _i32p, _dp, _dp, _dp]
        L.orc_apply_penalty.restype = C.c_double
        L.orc_apply_penalty.argtypes = [C.c_int64, _i64p, _i32p, _dp, _dp, _i64p, C.c_int64,
                                        C.c_double]
        L.orc_pcg_jacobi.restype = C.c_int
        L.orc_pcg_jacobi.argtypes = [C.c_int64, _i64p, _i32p, _dp, _dp, _dp, C.c_double,
                                     C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.orc_matfree_apply.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_sensitivities.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_double,
                                        _dp]
        L.orc_build_filter.restype = C.c_int64
        L.orc_build_filter.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
        L.orc_filter.argtypes = [C.c_int64, _i64p, _i64p, _dp, _dp, _dp, _dp]
        L.orc_oc_update.restype = C.c_double
        L.orc_oc_update.argtypes = [C.c_int64, _dp, _dp, _u8p, C.c_int64, C.c_double,
                                    C.c_double, C.c_double, C.POINTER(C.c_int),
                                    C.POINTER(C.c_double)]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.POINTER(OrcConfig), _dp, _dp, _i64p, C.c_int64, C.c_int,
                              C.POINTER(OrcIterResult), C.POINTER(C.c_int), C.c_void_p]
        L.orc_bracket_domain.restype = C.c_int64
        L.orc_bracket_domain.argtypes = [C.c_double, C.c_void_p, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_bracket_forces.restype = C.c_int
        L.orc_bracket_forces.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                         _dp]
        L.orc_bracket_fixed.restype = C.c_int64
        L.orc_bracket_fixed.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def ke0(E=1.0, nu=0.3, h=1.0):
    K = np.zeros(576)
    lib().orc_ke0(E, nu, h, K)
    return K.reshape(24, 24)


def element_dofs(i, j, k, nx, ny, nz):
    d = np.zeros(24, dtype=np.int64)
    lib().orc_element_dofs(i, j, k, nx, ny, nz, d)
    return d


def assemble_csr(nx, ny, nz, rho, h=1.0, E=1.0, nu=0.3, penal=3.0, Emin=1e-9, rho_min=0.01,
                 ref_compat=1):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    n = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    rp = np.zeros(n + 1, dtype=np.int64)
    nnz = lib().orc_assemble_csr(nx, ny, nz, h, E, nu, rho, penal, Emin, rho_min, ref_compat,
                                 rp.ctypes.data, None, None)
    col = np.zeros(nnz, dtype=np.int32)
    val = np.zeros(nnz)
    lib().orc_assemble_csr(nx, ny, nz, h, E, nu, rho, penal, Emin, rho_min, ref_compat,
                           rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    return rp, col, val


def spmv(rp, col, val, x):
    y = np.zeros(len(rp) - 1)
    lib().orc_csr_spmv(len(rp) - 1, rp, col, val, np.ascontiguousarray(x, np.float64), y)
    return y


def apply_penalty(rp, col, val, F, fixed, pf=1e8):
    fixed = np.ascontiguousarray(fixed, dtype=np.int64)
    return lib().orc_apply_penalty(len(rp) - 1, rp, col, val, F, fixed, len(fixed), pf)


def pcg(rp, col, val, b, tol=1e-8, maxit=10000):
    x = np.zeros(len(rp) - 1)
    it, err, tr = C.c_int(), C.c_double(), C.c_double()
    ok = lib().orc_pcg_jacobi(len(rp) - 1, rp, col, val, np.ascontiguousarray(b, np.float64),
                              x, tol, maxit, C.byref(it), C.byref(err), C.byref(tr))
    return x, dict(converged=bool(ok), iterations=it.value, rel_err=err.value,
                   true_rel=tr.value)


def matfree(nx, ny, nz, Ke, s_c, x):
    y = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    lib().orc_matfree_apply(nx, ny, nz, np.ascontiguousarray(Ke, np.float64).ravel(),
                            np.ascontiguousarray(s_c, np.float64),
                            np.ascontiguousarray(x, np.float64), y)
    return y


def sensitivities(nx, ny, nz, Ke, u, rho, penal=3.0):
    dc = np.zeros(nx * ny * nz)
    lib().orc_sensitivities(nx, ny, nz, np.ascontiguousarray(Ke, np.float64).ravel(),
                            np.ascontiguousarray(u, np.float64),
                            np.ascontiguousarray(rho, np.float64), penal, dc)
    return dc


def build_filter(nx, ny, nz, r):
    n_e = nx * ny * nz
    tot = lib().orc_build_filter(nx, ny, nz, r, None, None, None)
    offs = np.zeros(n_e + 1, dtype=np.int64)
    idx = np.zeros(tot, dtype=np.int64)
    w = np.zeros(tot)
    lib().orc_build_filter(nx, ny, nz, r, offs.ctypes.data, idx.ctypes.data, w.ctypes.data)
    return offs, idx, w


def apply_filter(offs, idx, w, dc, rho):
    out = np.zeros(len(offs) - 1)
    lib().orc_filter(len(offs) - 1, offs, idx, w, np.ascontiguousarray(dc, np.float64),
                     np.ascontiguousarray(rho, np.float64), out)
    return out


def oc_update(rho, dcf, design, vf, move, rho_min):
    rho = np.array(rho, dtype=np.float64)
    design = np.ascontiguousarray(design, dtype=np.uint8)
    npass, lm = C.c_int(), C.c_double()
    ch = lib().orc_oc_update(len(rho), rho, np.ascontiguousarray(dcf, np.float64), design,
                             int(design.sum()), vf, move, rho_min, C.byref(npass),
                             C.byref(lm))
    return rho, ch, npass.value, lm.value


def make_config(nx, ny, nz, **kw):
    d = dict(h=1.0, E=1.0, nu=0.3, volume_fraction=0.3, penalty=3.0, filter_radius=1.5,
             move_limit=0.2, rho_min=0.01, E_min=1e-9, max_iterations=50,
             convergence_tol=0.0, pcg_tol=1e-8, pcg_maxit=20000, penalty_factor=1e8,
             ref_compat=1, verbose=0)
    d.update(kw)
    return OrcConfig(nx=nx, ny=ny, nz=nz, **d)


def run(cfg, rho0, F, fixed, max_it=None):
    """Returns (final rho, list of per-iteration dicts, converged, u)."""
    max_it = cfg.max_iterations if max_it is None else max_it
    rho = np.array(rho0, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.float64)
    fixed = np.ascontiguousarray(fixed, dtype=np.int64)
    res = (OrcIterResult * max(max_it, 1))()
    conv = C.c_int()
    u = np.zeros(len(F))
    n = lib().orc_run(C.byref(cfg), rho, F, fixed, len(fixed), max_it, res, C.byref(conv),
                      u.ctypes.data)
    if n < 0:
        raise ValueError("oracle run failed (fixed DOF out of range)")
    hist = [{f: getattr(res[i], f) for f, _ in OrcIterResult._fields_} for i in range(n)]
    return rho, hist, bool(conv.value), u


def bracket(res_mm):
    nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
    n = lib().orc_bracket_domain(res_mm, None, C.byref(nx), C.byref(ny), C.byref(nz))
    g = np.zeros(n, dtype=np.int8)
    lib().orc_bracket_domain(res_mm, g.ctypes.data, C.byref(nx), C.byref(ny), C.byref(nz))
    return g, (nx.value, ny.value, nz.value)


def bracket_forces(nx, ny, nz, res_mm, load=15000.0):
    F = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    cnt = lib().orc_bracket_forces(nx, ny, nz, res_mm, load, F)
    return F, cnt


def bracket_fixed(nx, ny, nz):
    n = lib().orc_bracket_fixed(nx, ny, nz, None)
    f = np.zeros(n, dtype=np.int64)
    lib().orc_bracket_fixed(nx, ny, nz, f.ctypes.data)
    return f


def oc_tie_case():
    """One design element whose first bisection pass is an exact tie: lambda_1 = (1e-20 + 1e20)/2,
    dc = -4 (lambda_1 + 1e-20), so rho sqrt(-dc / (lambda_1 + 1e-20)) = 0.375 * 2 = 0.75 exactly
    (a power-of-two quotient and an exact square root) and vf = 0.75. The reference's strict test
    current_vf > vf_target (topology_optimizer.cpp:262-266) is false, so l2 = lambda_1 and every
    later pass evaluates a smaller lambda: the final density is above 0.75. Computing the
    candidate as rho sqrt(-dc) / sqrt(lambda + 1e-20) gives 0.7500000000000001 there, flips the
    first decision and ends below 0.75."""
    n = 16
    lam1 = 0.5 * (1e-20 + 1e20)
    rho = np.ones(n)                 # non-design (>= 0.99)
    dcf = -np.linspace(1.0, 2.0, n)
    rho[5] = 0.375
    dcf[5] = -4.0 * (lam1 * 1.0 + 1e-20)
    design = ((rho > 0.01 + 1e-6) & (rho < 0.99)).astype(np.uint8)
    return rho, dcf, design, 0.75, 0.5, 0.01  # rho, dcf, design, vf, move, rho_min
