"""Python mirror of the reference optimizer API over libamt.so.

Names follow the reference's C++ types (topology_optimizer.hpp:49-279,
element_kernel.hpp:61-87) and the Python binding the reference notebook imports
(notebooks/visualize_results.ipynb cell 23 lines 9-12: CPPTopologyOptimizer, SIMPConfig,
MaterialProperties, get_max_threads, is_available; run / run_iteration, cell 23-24).
Every call goes to the GPU through the C-ABI; nothing here computes on the CPU.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import _ffi


class AmtError(RuntimeError):
    pass


def _raise(code, msg):
    if code in (_ffi.AMT_EINVAL, _ffi.AMT_ENAN):
        raise ValueError(msg)  # std::invalid_argument
    if code == _ffi.AMT_ERANGE:
        raise IndexError(msg)  # std::out_of_range
    if code == _ffi.AMT_ENOTSUP:
        raise NotImplementedError(msg)
    raise AmtError(msg)


@dataclass
class MaterialProperties:
    """element_kernel.hpp:61-87 (Ti6Al4V defaults)."""
    E: float = 113.8e9
    nu: float = 0.342
    rho: float = 4430.0

    def lam(self):
        return self.E * self.nu / ((1.0 + self.nu) * (1.0 - 2.0 * self.nu))

    def mu(self):
        return self.E / (2.0 * (1.0 + self.nu))


@dataclass
class SIMPConfig:
    """topology_optimizer.hpp:49-59 plus the solver settings topology_optimizer.cpp:35-40
    fixes, and the device options."""
    volume_fraction: float = 0.25
    penalty: float = 3.0
    filter_radius: float = 2.0
    move_limit: float = 0.2
    rho_min: float = 0.01
    E_min: float = 1e-6
    max_iterations: int = 100
    convergence_tol: float = 0.01
    verbose: bool = True
    solver_tolerance: float = 1e-6
    solver_max_iterations: int = 5000
    penalty_factor: float = 1e8
    density_index: str = "reference"  # or "consistent"
    precision: str = "fp64"           # or "mixed_fp32", "mixed_bf16" (operator on MFMA)
    preconditioner: str = "jacobi"    # or "multigrid" (FP32 V-cycle inside FP64 flexible CG)
    warm_start: int = 0               # 1 (or True): previous displacement; 2: span of the last two
    inner_tolerance: float = 1e-4
    inner_max_iterations: int = 20000
    use_graphs: bool = True
    operator_tile: int = -1           # level-0 operator launch: -1 autotune, else a candidate
    mg_fused: bool = True             # fused V-cycle kernels on the small levels
    mg_coarse_cheb: int = 3           # Chebyshev degree on levels >= 2 (+1 on the two coarsest stencil levels of large grids; 0: Jacobi)


@dataclass
class IterationResult:
    """topology_optimizer.hpp:65-74 (+ solver detail)."""
    compliance: float = 0.0
    volume_fraction: float = 0.0
    max_density_change: float = 0.0
    solver_iterations: int = 0
    assembly_time_s: float = 0.0
    solve_time_s: float = 0.0
    sensitivity_time_s: float = 0.0
    update_time_s: float = 0.0
    solver_converged: bool = False
    outer_iterations: int = 0
    solver_fallback: bool = False
    true_rel_residual: float = 0.0
    bisect_passes: int = 0


@dataclass
class OptimizationResult:
    """topology_optimizer.hpp:80-87."""
    final_density: np.ndarray = None
    compliance_history: List[float] = field(default_factory=list)
    volume_history: List[float] = field(default_factory=list)
    iterations: int = 0
    converged: bool = False
    elapsed_time_s: float = 0.0
    iteration_results: List[IterationResult] = field(default_factory=list)


def _convert(r):
    return IterationResult(
        compliance=r.compliance, volume_fraction=r.volume_fraction,
        max_density_change=r.max_density_change, solver_iterations=r.solver_iterations,
        assembly_time_s=r.assembly_time_s, solve_time_s=r.solve_time_s,
        sensitivity_time_s=r.sensitivity_time_s, update_time_s=r.update_time_s,
        solver_converged=bool(r.solver_converged), outer_iterations=r.outer_iterations,
        solver_fallback=bool(r.solver_fallback), true_rel_residual=r.true_rel_residual,
        bisect_passes=r.bisect_passes)


def make_amt_config(nx, ny, nz, element_size, material, config):
    L = _ffi.load()
    c = _ffi.AmtConfig()
    L.amt_config_default(C.byref(c))
    c.nx, c.ny, c.nz = int(nx), int(ny), int(nz)
    c.element_size = float(element_size)
    c.E, c.nu = float(material.E), float(material.nu)
    c.volume_fraction = config.volume_fraction
    c.penalty = config.penalty
    c.filter_radius = config.filter_radius
    c.move_limit = config.move_limit
    c.rho_min = config.rho_min
    c.E_min = config.E_min
    c.max_iterations = int(config.max_iterations)
    c.convergence_tol = config.convergence_tol
    c.pcg_tol = config.solver_tolerance
    c.pcg_maxit = int(config.solver_max_iterations)
    c.penalty_factor = config.penalty_factor
    c.density_index = {"reference": _ffi.AMT_INDEX_REF_COMPAT,
                       "consistent": _ffi.AMT_INDEX_CONSISTENT}[config.density_index]
    c.precision = {"fp64": _ffi.AMT_PREC_FP64,
                   "mixed_fp32": _ffi.AMT_PREC_MIXED_FP32,
                   "mixed_bf16": _ffi.AMT_PREC_MIXED_BF16}[config.precision]
    c.precond = {"jacobi": _ffi.AMT_PRECOND_JACOBI, "unet": _ffi.AMT_PRECOND_UNET,
                 "multigrid": _ffi.AMT_PRECOND_MULTIGRID}[config.preconditioner]
    c.warm_start = int(config.warm_start)  # True -> 1; 2: combination of the last two solutions
    c.inner_tol = config.inner_tolerance
    c.inner_maxit = int(config.inner_max_iterations)
    c.verbose = 0
    c.use_graphs = int(bool(config.use_graphs))
    c.operator_tile = int(config.operator_tile)
    c.mg_fused = int(bool(config.mg_fused))
    c.mg_coarse_cheb = int(config.mg_coarse_cheb)
    return c


class TopologyOptimizer:
    """topology_optimizer.hpp:120-279 on one GPU."""

    def __init__(self, nx, ny, nz, element_size, material=None, config=None):
        self._L = _ffi.load()
        self.material = material or MaterialProperties()
        self.config = config or SIMPConfig()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.n_elements = self.nx * self.ny * self.nz
        self.n_dofs = 3 * (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self._h = C.c_void_p()
        cfg = make_amt_config(nx, ny, nz, element_size, self.material, self.config)
        rc = self._L.amt_create(C.byref(cfg), C.byref(self._h))
        if rc != _ffi.AMT_OK:
            _raise(rc, self._L.amt_last_error(None).decode())
        self._loads = None
        self._mask_built = False

    # -- plumbing --
    def _check(self, rc):
        if rc != _ffi.AMT_OK:
            _raise(rc, self._L.amt_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._L.amt_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_loads(self, forces, fixed_dofs):
        F = np.ascontiguousarray(forces, dtype=np.float64)
        fixed = np.ascontiguousarray(fixed_dofs, dtype=np.int64)
        key = (F.tobytes(), fixed.tobytes())
        if key == self._loads:
            return
        self._check(self._L.amt_set_loads(self._h, F, F.size, fixed, fixed.size))
        self._loads = key

    def set_density(self, rho):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        self._check(self._L.amt_set_density(self._h, rho, rho.size))

    def get_density(self):
        out = np.zeros(self.n_elements)
        self._check(self._L.amt_get_density(self._h, out, out.size))
        return out

    def get_displacement(self):
        out = np.zeros(self.n_dofs)
        self._check(self._L.amt_get_displacement(self._h, out, out.size))
        return out

    def begin_run(self):
        self._check(self._L.amt_begin_run(self._h))
        self._mask_built = True

    # -- reference API --
    def run_iteration(self, density, forces, fixed_dofs):
        """runIteration (topology_optimizer.cpp:294-384); density (numpy array) is updated
        in place when it is a float64 array, and returned."""
        self.set_loads(forces, fixed_dofs)
        self.set_density(density)
        if not self._mask_built:
            self.begin_run()
        r = _ffi.AmtIterationResult()
        self._check(self._L.amt_run_iteration(self._h, C.byref(r)))
        new = self.get_density()
        if isinstance(density, np.ndarray) and density.dtype == np.float64:
            density[...] = new.reshape(density.shape)
        elif isinstance(density, list):
            density[:] = new.tolist()
        return _convert(r)

    runIteration = run_iteration

    def step(self):
        """One OC iteration on the device-resident density (no host copies of the field)."""
        if not self._mask_built:
            self.begin_run()
        r = _ffi.AmtIterationResult()
        self._check(self._L.amt_run_iteration(self._h, C.byref(r)))
        return _convert(r)

    def run(self, initial_density, forces, fixed_dofs,
            callback: Optional[Callable] = None) -> OptimizationResult:
        """run (topology_optimizer.cpp:391-520)."""
        rho = np.asarray(initial_density, dtype=np.float64).ravel()
        if rho.size != self.n_elements:
            rho = np.full(self.n_elements, self.config.volume_fraction)
        self.set_loads(forces, fixed_dofs)
        self.set_density(rho)
        n = int(self.config.max_iterations)
        hist = (_ffi.AmtIterationResult * max(n, 1))()
        iters, conv, el = C.c_int(), C.c_int(), C.c_double()
        err = []

        def thunk(it, dptr, ne, rptr, user):
            try:
                d = np.ctypeslib.as_array(dptr, shape=(ne,)).copy()
                return 1 if callback(it, d, _convert(rptr.contents)) else 0
            except Exception as e:  # stop the run, re-raise below
                err.append(e)
                return 0

        cb = _ffi.ITER_CB(thunk) if callback else _ffi.ITER_CB()
        self._check(self._L.amt_run(self._h, n, hist, C.byref(iters), C.byref(conv),
                                    C.byref(el), cb, None))
        self._mask_built = True
        if err:
            raise err[0]
        res = [_convert(hist[i]) for i in range(iters.value)]
        return OptimizationResult(
            final_density=self.get_density(),
            compliance_history=[r.compliance for r in res],
            volume_history=[r.volume_fraction for r in res],
            iterations=iters.value, converged=bool(conv.value), elapsed_time_s=el.value,
            iteration_results=res)

    # -- phase entry points --
    def apply_operator(self, x, with_bc=True, precision="fp64"):
        """y = K x (penalty rows with with_bc) with the FP64 vector operator, or the FP32
        vector ("mixed_fp32") or split-BF16 MFMA ("mixed_bf16") one."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        prec = {"fp64": _ffi.AMT_PREC_FP64, "mixed_fp32": _ffi.AMT_PREC_MIXED_FP32,
                "mixed_bf16": _ffi.AMT_PREC_MIXED_BF16}[precision]
        self._check(self._L.amt_apply_operator_prec(self._h, x, y, int(bool(with_bc)), prec))
        return y

    def diagonal(self, with_bc=True):
        d = np.zeros(self.n_dofs)
        P = C.c_double()
        self._check(self._L.amt_get_diagonal(self._h, d, int(bool(with_bc)), C.byref(P)))
        return d, P.value

    def solve(self):
        u = np.zeros(self.n_dofs)
        st = _ffi.AmtSolveStats()
        self._check(self._L.amt_solve(self._h, u, C.byref(st)))
        return u, dict(iterations=st.iterations, converged=bool(st.converged),
                       outer_iterations=st.outer_iterations, fallback=bool(st.fallback),
                       rel_residual=st.rel_residual,
                       true_rel_residual=st.true_rel_residual, solve_time_s=st.solve_time_s)

    def apply_multigrid(self, r):
        """z = M r, one multigrid V-cycle (preconditioner="multigrid" only)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(self.n_dofs)
        self._check(self._L.amt_apply_multigrid(self._h, r, z))
        return z

    def sensitivities(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        dc = np.zeros(self.n_elements)
        self._check(self._L.amt_sensitivities(self._h, u, dc))
        return dc

    def filter(self, dc):
        dc = np.ascontiguousarray(dc, dtype=np.float64)
        out = np.zeros(self.n_elements)
        self._check(self._L.amt_filter(self._h, dc, out))
        return out

    def filter_neighbors(self, e):
        cnt = C.c_int()
        self._check(self._L.amt_filter_neighbors(self._h, int(e), None, None, 0, C.byref(cnt)))
        idx = np.zeros(max(cnt.value, 1), dtype=np.int64)
        w = np.zeros(max(cnt.value, 1))
        self._check(self._L.amt_filter_neighbors(self._h, int(e), idx.ctypes.data,
                                                 w.ctypes.data, cnt.value, C.byref(cnt)))
        return idx[:cnt.value], w[:cnt.value]

    def oc_update(self, dcf):
        dcf = np.ascontiguousarray(dcf, dtype=np.float64)
        ch, npass = C.c_double(), C.c_int()
        self._check(self._L.amt_oc_update(self._h, dcf, C.byref(ch), C.byref(npass)))
        return ch.value, npass.value

    def sync(self):
        self._check(self._L.amt_sync(self._h))

    # -- learned preconditioner --
    def attach_unet(self, net):
        self._check(self._L.amt_attach_unet(self._h, net._h if net is not None else None))
        self._net = net  # keep it alive while attached

    def collect_samples(self, max_samples):
        n = C.c_int()
        self._check(self._L.amt_unet_collect(self._h, int(max_samples), C.byref(n)))
        return n.value

    def collect_warm_samples(self, max_samples):
        """Arms warm-start sample collection (multigrid handle): later warm solves append their
        start residuals until max_samples are held. Returns the samples held now."""
        n = C.c_int()
        self._check(self._L.amt_unet_collect_warm(self._h, int(max_samples), C.byref(n)))
        return n.value

    def train_unet(self, net, steps, lr=1e-3, with_jacobi=False):
        """Adam steps on the collected samples; returns each step's energy loss (and, with
        with_jacobi, the loss of the Jacobi term alone on the same sample)."""
        losses = np.zeros(max(int(steps), 1))
        jl = np.zeros_like(losses)
        self._check(self._L.amt_unet_train(self._h, net._h, int(steps), float(lr),
                                           losses.ctypes.data, jl.ctypes.data))
        n = int(steps)
        return (losses[:n], jl[:n]) if with_jacobi else losses[:n]

    def profile_enable(self, on=True):
        self._check(self._L.amt_profile_enable(self._h, int(bool(on))))

    def profile_read(self, kind=0):
        """(timed launches, their total ms) of kind 0 (operator) or 1 (level-0 pre-smoothing
        with the flexible-CG update)."""
        n, ms = C.c_int64(), C.c_double()
        self._check(self._L.amt_profile_read_kind(self._h, int(kind), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def operator_info(self, kind=0):
        """Level-0 operator launch of kind 0 (FP64 CG operator), 1 (FP32 pre-smoothing with the
        flexible-CG update) or 2 (FP32 post-smoothing): dict(candidate, tj, tk, planes, us)."""
        v = [C.c_int() for _ in range(4)]
        us = C.c_double()
        self._check(self._L.amt_operator_info(self._h, int(kind), *[C.byref(x) for x in v],
                                              C.byref(us)))
        return dict(candidate=v[0].value, tj=v[1].value, tk=v[2].value, planes=v[3].value,
                    us=us.value)

    def operator_timings(self, kind=0):
        """The autotuner's launch time (microseconds) of every candidate for that kind."""
        out = np.zeros(16)
        self._check(self._L.amt_operator_timings(self._h, int(kind), out, out.size))
        return out


def element_stiffness(E=1.0, nu=0.3, h=1.0):
    """ElementKernel::computeStiffness (element_kernel.cpp:155-179) -> (24, 24) array."""
    K = np.zeros(576)
    rc = _ffi.load().amt_element_stiffness(float(E), float(nu), float(h), K)
    if rc != _ffi.AMT_OK:
        _raise(rc, _ffi.load().amt_last_error(None).decode())
    return K.reshape(24, 24)


class UNet:
    """BF16 3D U-Net of the learned preconditioner (4 levels, 16/32/64/128 channels,
    3x3x3 convolutions) on a padded voxel grid; tensors are [px*py*pz, 16] channels-last."""

    TENSORS = ("X0", "A0a", "S0", "P1", "A1a", "S1", "P2", "A2a", "S2", "P3", "B3a", "B3", "U2",
               "C2a", "C2", "U1", "C1a", "C1", "U0", "C0a", "C0", "OUT")
    TENSOR_LEVEL_CHANNELS = {
        "X0": (0, 16), "A0a": (0, 16), "S0": (0, 16), "P1": (1, 16), "A1a": (1, 32),
        "S1": (1, 32), "P2": (2, 32), "A2a": (2, 64), "S2": (2, 64), "P3": (3, 64),
        "B3a": (3, 128), "B3": (3, 128), "U2": (2, 128), "C2a": (2, 64), "C2": (2, 64),
        "U1": (1, 64), "C1a": (1, 32), "C1": (1, 32), "U0": (0, 32), "C0a": (0, 16),
        "C0": (0, 16), "OUT": (0, 16)}

    def __init__(self, px=None, py=None, pz=None, seed=1, optimizer=None):
        self._L = _ffi.load()
        self._h = C.c_void_p()
        if optimizer is not None:
            rc = self._L.amt_unet_create_for(optimizer._h, int(seed), C.byref(self._h))
            px, py, pz = (-(-(n + 1) // 8) * 8 for n in (optimizer.nx, optimizer.ny, optimizer.nz))
        else:
            rc = self._L.amt_unet_create(int(px), int(py), int(pz), int(seed), C.byref(self._h))
        if rc != _ffi.AMT_OK:
            _raise(rc, self._L.amt_last_error(None).decode())
        self.n_params = int(self._L.amt_unet_num_params(self._h))
        self.shape = (int(px), int(py), int(pz))

    def _check(self, rc):
        if rc != _ffi.AMT_OK:
            _raise(rc, self._L.amt_last_error(None).decode())

    def close(self):
        if self._h:
            self._L.amt_unet_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        self._check(self._L.amt_unet_get_params(self._h, out, out.size))
        return out

    def set_params(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self._L.amt_unet_set_params(self._h, params, params.size))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros_like(x)
        self._check(self._L.amt_unet_forward(self._h, x, y))
        return y

    def infer(self, x):
        """The inference forward pass (the same output as forward, no training state)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros_like(x)
        self._check(self._L.amt_unet_infer(self._h, x, y))
        return y

    def backward(self, dy):
        dy = np.ascontiguousarray(dy, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        self._check(self._L.amt_unet_backward(self._h, dy, g))
        return g

    @property
    def gain(self):
        g = C.c_double()
        self._check(self._L.amt_unet_get_gain(self._h, C.byref(g)))
        return g.value

    @gain.setter
    def gain(self, value):
        self._check(self._L.amt_unet_set_gain(self._h, float(value)))

    def benchmark(self, reps=5):
        """(milliseconds per forward pass, convolution TFLOP/s) on the stored input."""
        ms, tf = C.c_double(), C.c_double()
        self._check(self._L.amt_unet_benchmark(self._h, int(reps), C.byref(ms), C.byref(tf)))
        return ms.value, tf.value

    def activation(self, name):
        """Stored activation of the last forward, shape (px>>l, py>>l, pz>>l, channels)."""
        i = self.TENSORS.index(name)
        lev, ch = self.TENSOR_LEVEL_CHANNELS[name]
        shape = tuple(n >> lev for n in self.shape) + (ch,)
        out = np.zeros(shape, dtype=np.float32)
        self._check(self._L.amt_unet_activation(self._h, i, out, out.size))
        return out


class BatchOptimizer:
    """Many independent problems optimised on one GPU (amt_batch_*). problems: list of
    (nx, ny, nz, element_size, MaterialProperties, SIMPConfig); precision, warm_start and the
    solver must agree; density index consistent.
      - Jacobi-PCG: all problems in lockstep, every phase one persistent launch over all of
        them, the OC step replayed from hipGraphs;
      - preconditioner="multigrid": one single-problem handle and stream per problem, run by
        eight host threads so the problems' launches overlap on the device."""

    def __init__(self, problems):
        self._L = _ffi.load()
        cfgs = (_ffi.AmtConfig * len(problems))()
        for t, (nx, ny, nz, h, mat, cfg) in enumerate(problems):
            cfgs[t] = make_amt_config(nx, ny, nz, h, mat or MaterialProperties(),
                                      cfg or SIMPConfig())
        self._h = C.c_void_p()
        rc = self._L.amt_batch_create(C.cast(cfgs, C.c_void_p), len(problems), C.byref(self._h))
        if rc != _ffi.AMT_OK:
            _raise(rc, self._L.amt_last_error(None).decode())
        self.n = len(problems)
        self.sizes = []
        for t in range(self.n):
            ne, nd = C.c_int64(), C.c_int64()
            self._check(self._L.amt_batch_size(self._h, t, C.byref(ne), C.byref(nd)))
            self.sizes.append((ne.value, nd.value))

    def _check(self, rc):
        if rc != _ffi.AMT_OK:
            _raise(rc, self._L.amt_last_error(None).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.amt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_loads(self, i, forces, fixed_dofs):
        F = np.ascontiguousarray(forces, dtype=np.float64)
        fx = np.ascontiguousarray(fixed_dofs, dtype=np.int64)
        self._check(self._L.amt_batch_set_loads(self._h, int(i), F, F.size, fx, fx.size))

    def set_density(self, i, rho):
        r = np.ascontiguousarray(rho, dtype=np.float64)
        self._check(self._L.amt_batch_set_density(self._h, int(i), r, r.size))

    def get_density(self, i):
        out = np.zeros(self.sizes[i][0])
        self._check(self._L.amt_batch_get_density(self._h, int(i), out, out.size))
        return out

    def get_displacement(self, i):
        out = np.zeros(self.sizes[i][1])
        self._check(self._L.amt_batch_get_displacement(self._h, int(i), out, out.size))
        return out

    def run(self, iterations):
        """Runs OC iterations; returns (history[iteration][problem] of IterationResult,
        elapsed seconds)."""
        hist = (_ffi.AmtIterationResult * max(int(iterations) * self.n, 1))()
        el = C.c_double()
        self._check(self._L.amt_batch_run(self._h, int(iterations), C.cast(hist, C.c_void_p),
                                          C.byref(el)))
        return ([[_convert(hist[it * self.n + t]) for t in range(self.n)]
                 for it in range(int(iterations))], el.value)

    # lockstep matvec candidates (amt_batch_operator_info): 0 the gather tile, 1 + 3 s + l the
    # streaming tile of shape BATCH_SHAPES[s] with BATCH_PLANES[l] node planes
    BATCH_SHAPES = ((16, 16), (8, 32), (4, 64))
    BATCH_PLANES = (2, 4, 8)

    def operator_info(self):
        """(chosen matvec candidate, {candidate: measured launch time in us}) of a lockstep
        batch; the candidate is -1 before the first run."""
        c = C.c_int()
        us = np.zeros(16)
        self._check(self._L.amt_batch_operator_info(self._h, C.byref(c), us, us.size))
        return c.value, {i: float(t) for i, t in enumerate(us) if t > 0}

    @classmethod
    def operator_shape(cls, cand):
        """(tile description, node planes) of lockstep matvec candidate cand."""
        if cand == 0:
            return "4x4x16 node gather tile", 0
        s, l = divmod(cand - 1, len(cls.BATCH_PLANES))
        return "%dx%d streaming tile" % cls.BATCH_SHAPES[s], cls.BATCH_PLANES[l]


# notebook binding names (cell 23 lines 9-12)
CPPTopologyOptimizer = TopologyOptimizer


def is_available():
    """True when libamt.so loads and a HIP device is present."""
    try:
        L = _ffi.load()
    except Exception:
        return False
    n = C.c_int()
    return L.amt_device_info(None, 0, C.byref(n)) == _ffi.AMT_OK


def get_max_threads():
    """The notebook prints the engine's parallel width; here: the GPU's compute units."""
    n = C.c_int()
    if _ffi.load().amt_device_info(None, 0, C.byref(n)) != _ffi.AMT_OK:
        return 0
    return n.value


def device_name():
    buf = C.create_string_buffer(256)
    n = C.c_int()
    if _ffi.load().amt_device_info(buf, 256, C.byref(n)) != _ffi.AMT_OK:
        return None
    return buf.value.decode()
