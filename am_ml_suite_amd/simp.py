"""The suite's Python SIMP loop (src/am/numerical/topopt.py:84-411) on the MI355X kernels.

Same constructor, parameters, run loop and result as the reference's SIMPOptimizer; every
phase of an iteration (density filter, stiffness scales, per-load-case solve with constrained
and orphan DOFs eliminated, compliance and sensitivities, sensitivity filter, OC bisection)
runs in libamt.so (include/amt.h, amt_simp_*). The reference solves K_ff u_f = F_f with a
sparse direct solver (fem.py:301-304); here FP64 Jacobi-PCG runs to `solve_tol` (relative
residual) and each iteration reports the FP64 true residual.
"""
import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import _ffi
from .optimizer import _raise


@dataclass
class MaterialProperties:
    """fem.py:17-41 (Ti6Al4V defaults, SI units)."""
    E: float = 113.8e9
    nu: float = 0.342
    rho: float = 4430.0

    @property
    def lame_lambda(self):
        return self.E * self.nu / ((1 + self.nu) * (1 - 2 * self.nu))

    @property
    def lame_mu(self):
        return self.E / (2 * (1 + self.nu))


@dataclass
class SIMPParams:
    """topopt.py:30-54, plus the GPU solve's stopping rule."""
    volume_fraction: float = 0.25
    penalty: float = 3.0
    filter_radius: float = 2.0
    move_limit: float = 0.2
    rho_min: float = 0.01
    max_iterations: int = 100
    convergence_tol: float = 0.01
    E_min: float = 1e-9
    solve_tol: float = 1e-10
    solve_maxit: int = 200000


@dataclass
class TOResult:
    """topopt.py:57-81."""
    density: np.ndarray
    compliance_history: List[float] = field(default_factory=list)
    volume_history: List[float] = field(default_factory=list)
    iterations: int = 0
    converged: bool = False
    elapsed_time: float = 0.0
    solver_iterations: List[int] = field(default_factory=list)
    true_rel_residual: List[float] = field(default_factory=list)

    @property
    def final_compliance(self):
        return self.compliance_history[-1] if self.compliance_history else 0.0

    @property
    def final_volume(self):
        return self.volume_history[-1] if self.volume_history else 0.0


class SIMPOptimizer:
    """topopt.py:84-411. domain: VoxelDomain (grid (nx, ny, nz) int8, resolution_mm);
    load_cases: objects with get_force_vector(n_dofs) and get_constrained_dofs()."""

    def __init__(self, domain, load_cases, material: MaterialProperties = None,
                 params: SIMPParams = None, verbose: bool = True):
        self.domain = domain
        self.load_cases = list(load_cases)
        self.material = material or MaterialProperties()
        self.params = params or SIMPParams()
        self.verbose = verbose
        self.nx, self.ny, self.nz = domain.shape
        self.n_elements = self.nx * self.ny * self.nz
        self.element_size = domain.resolution_mm / 1000  # mm -> m (topopt.py:121)
        self.n_dofs = 3 * (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self.design_mask = np.asarray(domain.grid) == 0
        self._L = _ffi.load()
        p = self.params
        prm = _ffi.AmtSimpParams(p.volume_fraction, p.penalty, p.filter_radius, p.move_limit,
                                 p.rho_min, p.convergence_tol, p.E_min, p.max_iterations,
                                 p.solve_tol, p.solve_maxit)
        grid = np.ascontiguousarray(np.asarray(domain.grid, dtype=np.int8))
        h = C.c_void_p()
        rc = self._L.amt_simp_create(self.nx, self.ny, self.nz, self.element_size,
                                     self.material.E, self.material.nu, grid, C.byref(prm),
                                     C.byref(h))
        if rc:
            _raise(rc, self._L.amt_last_error(None).decode())
        self._h = h
        for lc in self.load_cases:
            F = np.ascontiguousarray(lc.get_force_vector(self.n_dofs), dtype=np.float64)
            cons = np.ascontiguousarray(lc.get_constrained_dofs(), dtype=np.int64)
            self._check(self._L.amt_simp_add_load_case(h, F, F.size, cons, cons.size))
        # topopt.py:124: uniform start at the volume fraction
        self.density = np.full((self.nx, self.ny, self.nz), p.volume_fraction)

    def _check(self, rc):
        if rc:
            _raise(rc, self._L.amt_last_error(None).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.amt_simp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply_filter(self, x):
        """ndimage.convolve(x, kernel, mode='reflect') (topopt.py:153-156)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.nx, self.ny, self.nz)
        y = np.empty_like(x)
        self._check(self._L.amt_simp_filter(self._h, x.ravel(), y.ravel(), x.size))
        return y

    def displacement(self, lc=0):
        u = np.empty(self.n_dofs)
        self._check(self._L.amt_simp_get_displacement(self._h, lc, u, u.size))
        return u

    def step(self, density):
        """One pass of run()'s loop body from `density`; returns (new density, result)."""
        d = np.ascontiguousarray(density, dtype=np.float64).ravel()
        self._check(self._L.amt_simp_set_density(self._h, d, d.size))
        it = _ffi.AmtSimpIteration()
        self._check(self._L.amt_simp_step(self._h, C.byref(it)))
        out = np.empty(self.n_elements)
        self._check(self._L.amt_simp_get_density(self._h, out, out.size))
        return out.reshape(self.nx, self.ny, self.nz), it

    def run(self, callback: Optional[Callable] = None) -> TOResult:
        """topopt.py:264-383."""
        t0 = time.time()
        res = TOResult(density=self.density.copy())
        density = self.density.copy()
        if self.verbose:
            print("SIMP Topology Optimization (MI355X)")
            print(f"Grid size: {self.nx} x {self.ny} x {self.nz}")
            print(f"Design elements: {int(self.design_mask.sum()):,}")
        for iteration in range(self.params.max_iterations):
            density, it = self.step(density)
            res.compliance_history.append(it.compliance)
            res.volume_history.append(it.volume)
            res.solver_iterations.append(it.solver_iterations)
            res.true_rel_residual.append(it.max_true_rel_residual)
            if self.verbose:
                print(f"Iter {iteration+1:3d} | Compliance: {it.compliance:.4e} | "
                      f"Volume: {it.volume:.4f} | Change: {it.change:.4f}")
            if callback:
                callback(iteration, density, it.compliance)
            res.iterations = iteration + 1
            if it.change < self.params.convergence_tol:
                res.converged = True
                break
        res.density = density
        res.elapsed_time = time.time() - t0
        return res


def threshold_density(density, threshold=0.5):
    """topopt.py:386-402."""
    return (np.asarray(density) >= threshold).astype(np.float64)
