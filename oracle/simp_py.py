"""CPU oracle of the suite's Python SIMP loop (src/am/numerical/topopt.py:84-411 with
src/am/numerical/fem.py:44-384), restated with numpy and scipy (the reference's own
dependencies: scipy.ndimage.convolve for the filters, scipy.sparse + spsolve for the solve).

Test infrastructure: only tests/ and bench.py's cpu_baseline leg use it, as the checker and
the CPU baseline. The product (am_ml_suite_amd.simp) never imports it.

Pinned by tests/test_simp_oracle.py: the filter against a brute-force convolution with the
'reflect' extension written out; Ke0 against its closed-form entries and rigid-body null
space; the solve by its residual and the Galerkin identity c = F.u; the sensitivities by
finite differences of the compliance; the OC update by its bounds, move limit and volume
bracket.
"""
import numpy as np
from scipy import ndimage, sparse
from scipy.sparse.linalg import spsolve

NODE_OFFSETS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]  # fem.py:145-148


def element_stiffness(h, E, nu):
    """fem.py:44-130: 2x2x2 Gauss integration of B^T C B over a cube of edge h."""
    a = h / 2
    C = E / ((1 + nu) * (1 - 2 * nu)) * np.array([
        [1 - nu, nu, nu, 0, 0, 0], [nu, 1 - nu, nu, 0, 0, 0], [nu, nu, 1 - nu, 0, 0, 0],
        [0, 0, 0, (1 - 2 * nu) / 2, 0, 0], [0, 0, 0, 0, (1 - 2 * nu) / 2, 0],
        [0, 0, 0, 0, 0, (1 - 2 * nu) / 2]])
    nodes = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
    g = 1 / np.sqrt(3)
    Ke = np.zeros((24, 24))
    for xi, eta, zeta in nodes * g:
        dN = np.empty((8, 3))
        dN[:, 0] = 0.125 * nodes[:, 0] * (1 + nodes[:, 1] * eta) * (1 + nodes[:, 2] * zeta)
        dN[:, 1] = 0.125 * nodes[:, 1] * (1 + nodes[:, 0] * xi) * (1 + nodes[:, 2] * zeta)
        dN[:, 2] = 0.125 * nodes[:, 2] * (1 + nodes[:, 0] * xi) * (1 + nodes[:, 1] * eta)
        dx = dN / a
        B = np.zeros((6, 24))
        B[0, 0::3] = dx[:, 0]
        B[1, 1::3] = dx[:, 1]
        B[2, 2::3] = dx[:, 2]
        B[3, 0::3], B[3, 1::3] = dx[:, 1], dx[:, 0]
        B[4, 1::3], B[4, 2::3] = dx[:, 2], dx[:, 1]
        B[5, 0::3], B[5, 2::3] = dx[:, 2], dx[:, 0]
        Ke += B.T @ C @ B * a ** 3
    return Ke


def element_dofs(nx, ny, nz):
    """fem.py:133-165 for every element, C order: (n_e, 24) global DOFs."""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cols = []
    for di, dj, dk in NODE_OFFSETS:
        node = (i + di) * (ny + 1) * (nz + 1) + (j + dj) * (nz + 1) + (k + dk)
        cols += [3 * node, 3 * node + 1, 3 * node + 2]
    return np.stack([c.ravel() for c in cols], axis=1).astype(np.int64)


def cone_kernel(r):
    """topopt.py:134-151."""
    size = int(2 * np.ceil(r) + 1)
    c = size // 2
    i, j, k = np.meshgrid(np.arange(size), np.arange(size), np.arange(size), indexing="ij")
    d = np.sqrt((i - c) ** 2 + (j - c) ** 2 + (k - c) ** 2)
    w = np.where(d <= r, r - d, 0.0)
    return w / w.sum()


def apply_filter(x, w):
    """topopt.py:153-156."""
    return ndimage.convolve(x, w, mode="reflect")


def scales(grid, rhof, E, E_min, p):
    """fem.py:181-207: E_eff / E per element (void 0)."""
    return np.where(grid == -1, 0.0, np.where(grid == 1, 1.0, (E_min + rhof ** p * (E - E_min)) / E))


def assemble(grid, rhof, Ke, E, E_min, p):
    """fem.py:168-222 (vectorised): K = sum_e s_e Ke0, then (K + K^T) / 2."""
    nx, ny, nz = grid.shape
    s = scales(grid, rhof, E, E_min, p).ravel()
    edofs = element_dofs(nx, ny, nz)
    keep = grid.ravel() != -1
    rows = np.repeat(edofs[keep], 24, axis=1).ravel()
    cols = np.tile(edofs[keep], (1, 24)).ravel()
    data = (s[keep][:, None, None] * Ke[None]).ravel()
    n = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    K = sparse.coo_matrix((data, (rows, cols)), shape=(n, n)).tocsr()
    return (K + K.T) / 2


def solve(K, F, constrained):
    """fem.py:225-307 (direct): orphan DOFs (|K_ii| < 1e-15) join the constrained ones."""
    n = K.shape[0]
    orphans = np.where(np.abs(K.diagonal()) < 1e-15)[0]
    eff = np.unique(np.concatenate([np.asarray(constrained, dtype=np.int64), orphans]))
    free = np.setdiff1d(np.arange(n), eff)
    u = np.zeros(n)
    if free.size:
        Kff = K[free, :][:, free]
        u[free] = spsolve(Kff.tocsc(), F[free])
    return u, free


def element_energy(u, Ke, nx, ny, nz):
    """u_e^T Ke0 u_e per element (C order)."""
    ue = u[element_dofs(nx, ny, nz)]
    return np.einsum("ei,ij,ej->e", ue, Ke, ue)


def compliance(u, grid, rhof, Ke, E, E_min, p):
    """fem.py:310-384."""
    nx, ny, nz = grid.shape
    ce = element_energy(u, Ke, nx, ny, nz).reshape(grid.shape)
    ratio = np.where(grid == 0, E_min / E + rhof ** p * (1 - E_min / E), 1.0)
    return float(np.sum(np.where(grid == -1, 0.0, ratio * ce))), ce


def sensitivities(ce, grid, rhof, p):
    """topopt.py:158-198."""
    return np.where(grid == 0, -p * rhof ** (p - 1) * ce, 0.0)


def oc_update(density, dc, design, vf, move, rho_min):
    """topopt.py:200-262. Returns (rho_new, passes)."""
    l1, l2 = 1e-9, 1e9
    n_design = np.sum(design)
    rho_new, passes = density, 0
    for _ in range(100):
        if l1 + l2 < 1e-12:
            break
        if (l2 - l1) / max((l1 + l2), 1e-10) <= 1e-3:
            break
        lmid = 0.5 * (l1 + l2)
        Be = np.sqrt(-dc / (1.0 * lmid + 1e-10))
        rho_new = np.maximum(rho_min, np.maximum(density - move,
                                                 np.minimum(1.0, np.minimum(density + move,
                                                                            density * Be))))
        rho_new = np.where(design, rho_new, density)
        passes += 1
        if np.sum(rho_new[design]) / n_design > vf:
            l1 = lmid
        else:
            l2 = lmid
    return rho_new, passes


class Problem:
    """What SIMPOptimizer.__init__ sets up (topopt.py:97-131)."""

    def __init__(self, grid, resolution_mm, loads, E=113.8e9, nu=0.342, volume_fraction=0.25,
                 penalty=3.0, filter_radius=2.0, move_limit=0.2, rho_min=0.01,
                 convergence_tol=0.01, E_min=1e-9):
        """loads: list of (F, constrained DOFs)."""
        self.grid = np.asarray(grid, dtype=np.int8)
        self.h = resolution_mm / 1000
        self.loads = loads
        self.E, self.nu = E, nu
        self.vf, self.p, self.r = volume_fraction, penalty, filter_radius
        self.move, self.rho_min, self.tol, self.E_min = move_limit, rho_min, convergence_tol, E_min
        self.Ke = element_stiffness(self.h, E, nu)
        self.w = cone_kernel(filter_radius)
        self.design = self.grid == 0


def step(P, density):
    """One pass of run()'s loop body (topopt.py:298-351). Returns (density_new, info)."""
    rhof = np.clip(apply_filter(density, P.w), P.rho_min, 1.0)
    K = assemble(P.grid, rhof, P.Ke, P.E, P.E_min, P.p)
    total, dc_total, us = 0.0, np.zeros_like(density), []
    for F, cons in P.loads:
        u, _ = solve(K, F, cons)
        c, ce = compliance(u, P.grid, rhof, P.Ke, P.E, P.E_min, P.p)
        total += c
        dc_total += sensitivities(ce, P.grid, rhof, P.p)
        us.append(u)
    dcf = apply_filter(dc_total * density, P.w) / np.maximum(density, 1e-3)
    new, passes = oc_update(density, dcf, P.design, P.vf, P.move, P.rho_min)
    return new, dict(compliance=total, volume=float(np.mean(new[P.design])),
                     change=float(np.max(np.abs(new - density))), u=us, rhof=rhof, dcf=dcf,
                     passes=passes)


def run(P, max_iterations):
    """topopt.py:264-383 without printing."""
    density = np.full(P.grid.shape, P.vf)
    hist = []
    for _ in range(max_iterations):
        density, info = step(P, density)
        hist.append(info)
        if info["change"] < P.tol:
            break
    return density, hist
