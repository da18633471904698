"""Problem definitions: the inputs of the optimizer (density field, forces, fixed DOFs).

- VoxelDomain / LoadCase mirror the reference's interface types (src/am/core/geometry.py:16-76,
  src/am/core/loads.py:107-156): an int8 grid (-1 void, 0 design, 1 fixed) and loads/BCs
  that turn into F[n_dof] and a fixed-DOF list.
- bracket_domain / bracket_load_case restate the BRK-A-01 set-up of the reference's C++
  driver (cpp_engine/src/am_topopt.cpp:98-334, initial density :561-572).
- cantilever_problem / bracket_like_problem generate the synthetic benchmark problems
  (30x10x10 cantilever; 120x60x60 bracket-like domain with solid pads, a void box and seeded
  point loads) from a splitmix64 stream, so every run sees the same arrays.
- batch_problems generates the dataset-generation set: many problems of uneven sizes, each
  with one clamped face and 1-4 point loads on the other faces.
"""
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

VOID, DESIGN, FIXED = -1, 0, 1


@dataclass
class VoxelDomain:
    """geometry.py:16-76. grid has shape (nx, ny, nz), indexed [i, j, k]."""
    grid: np.ndarray
    resolution_mm: float
    origin: Tuple[float, float, float] = (0.0, 0.0, 0.0)

    @property
    def shape(self):
        return tuple(self.grid.shape)

    @property
    def n_elements(self):
        return int(np.prod(self.grid.shape))

    @property
    def n_design(self):
        return int(np.sum(self.grid == DESIGN))

    @property
    def n_fixed(self):
        return int(np.sum(self.grid == FIXED))

    @property
    def n_void(self):
        return int(np.sum(self.grid == VOID))

    def slot_order(self, values):
        """(nx,ny,nz) array -> flat vector in the optimizer's x-fastest slot order."""
        return np.ascontiguousarray(np.asarray(values).transpose(2, 1, 0)).ravel()

    def c_order(self, values):
        """(nx,ny,nz) array -> flat C-order vector (i*ny*nz + j*nz + k)."""
        return np.ascontiguousarray(values).ravel()

    def initial_density(self, volume_fraction, void_value, order="slot"):
        """am_topopt.cpp:561-572: design=vf, fixed=1, void=void_value."""
        g = self.grid
        rho = np.where(g == DESIGN, volume_fraction, np.where(g == FIXED, 1.0, void_value))
        return self.slot_order(rho) if order == "slot" else self.c_order(rho)


@dataclass
class LoadCase:
    """loads.py:107-156 reduced to what the optimizer consumes: point forces per node and
    constrained DOFs."""
    name: str
    n_dofs: int
    forces: List[Tuple[int, Tuple[float, float, float]]] = field(default_factory=list)
    constrained: List[int] = field(default_factory=list)

    def get_force_vector(self, n_dofs=None):
        F = np.zeros(self.n_dofs if n_dofs is None else n_dofs)
        for node, f in self.forces:
            for c in range(3):
                F[3 * node + c] += f[c]
        return F

    def get_constrained_dofs(self):
        return np.unique(np.asarray(self.constrained, dtype=np.int64))


def node_id(i, j, k, ny, nz):
    """Node index, z fastest (element_kernel.cpp:207)."""
    return (i * (ny + 1) + j) * (nz + 1) + k


# ----------------------------------------------------------------------------------------
# BRK-A-01 (am_topopt.cpp defaults: 120x60x80 mm box)
# ----------------------------------------------------------------------------------------
def bracket_domain(resolution_mm):
    """BracketDomain (am_topopt.cpp:98-239) as a VoxelDomain (grid[i, j, k])."""
    res = float(resolution_mm)
    nx, ny, nz = int(120.0 / res), int(60.0 / res), int(80.0 / res)
    g = np.zeros((nx, ny, nz), dtype=np.int8)

    def cyl(cx, cy, z0, rad, hgt, val):
        i0, i1 = max(0, int(cx - rad - 1)), min(nx, int(cx + rad + 2))
        j0, j1 = max(0, int(cy - rad - 1)), min(ny, int(cy + rad + 2))
        for i in range(i0, i1):
            for j in range(j0, j1):
                di, dj = i - cx, j - cy
                if di * di + dj * dj <= rad * rad:
                    for k in range(z0, min(nz, z0 + hgt)):
                        if val == VOID or g[i, j, k] != VOID:
                            g[i, j, k] = val

    base_z = max(1, int(5.0 / res))
    g[:, :, :min(base_z, nz)] = FIXED
    hole_r = 8.0 / 2.0 / res
    hox, hoy = int(15.0 / res), int(15.0 / res)
    for hx, hy in ((hox, hoy), (nx - hox, hoy), (hox, ny - hoy), (nx - hox, ny - hoy)):
        cyl(hx, hy, 0, hole_r, base_z, VOID)
    ezo, eh = int(15.0 / res), max(1, int(20.0 / res))
    ezc, ex, ey = nz - ezo, nx // 2, ny // 2
    z0 = max(0, ezc - eh // 2)
    z1 = min(nz, z0 + eh)
    if z1 - z0 > 0:
        cyl(ex, ey, z0, (12.0 / 2.0 + 4.0) / res, z1 - z0, FIXED)
        cyl(ex, ey, z0, 12.0 / 2.0 / res, z1 - z0, VOID)
    if z0 > base_z:
        cyl(ex, ey, base_z, (12.0 / 2.0 + 4.0) / res, z0 - base_z, FIXED)
    return VoxelDomain(grid=g, resolution_mm=res)


def bracket_load_case(domain, load_N=15000.0):
    """createForceVector + createFixedDofs (am_topopt.cpp:248-334): a -z ring load on the
    eyelet and the z=0 clamp."""
    nx, ny, nz = domain.shape
    res = domain.resolution_mm
    n_dofs = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    ezo, eh = int(15.0 / res), max(1, int(20.0 / res))
    ezc = nz - ezo
    z0 = max(0, ezc - eh // 2)
    z1 = min(nz, z0 + eh)
    ez = max(1, min(nz - 1, (z0 + z1) // 2))
    ex, ey = nx // 2, ny // 2
    inner_mm = 12.0 / 2.0 + res * 1.25
    outer_mm = 12.0 / 2.0 + 4.0
    if inner_mm >= outer_mm:
        inner_mm = 12.0 / 2.0 + 4.0 * 0.5
    ri, ro = inner_mm / res, outer_mm / res
    search = int(ro) + 2
    nodes = []
    for i in range(max(0, ex - search), min(nx, ex + search) + 1):
        for j in range(max(0, ey - search), min(ny, ey + search) + 1):
            d = np.sqrt(float(i - ex) ** 2 + float(j - ey) ** 2)
            if ri < d <= ro:
                nodes.append(node_id(i, j, ez, ny, nz))
    if not nodes:
        nodes.append(node_id(ex, ey, ez, ny, nz))
    fpn = load_N / len(nodes)
    lc = LoadCase("BRK-A-01_Static_1", n_dofs)
    lc.forces = [(n, (0.0, 0.0, -fpn)) for n in nodes]
    lc.constrained = [3 * node_id(i, j, 0, ny, nz) + c
                      for i in range(nx + 1) for j in range(ny + 1) for c in range(3)]
    return lc


# ----------------------------------------------------------------------------------------
# Synthetic benchmark problems
# ----------------------------------------------------------------------------------------
class SplitMix64:
    """splitmix64 stream (Steele, Lea, Flood 2014)."""
    M = (1 << 64) - 1

    def __init__(self, seed):
        self.s = seed & self.M

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.M
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.M
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def randint(self, a, b):
        """uniform integer in [a, b]"""
        return a + self.next() % (b - a + 1)

    def unit_vector(self):
        z = self.uniform(-1.0, 1.0)
        phi = self.uniform(0.0, 2.0 * np.pi)
        s = np.sqrt(max(0.0, 1.0 - z * z))
        return (s * np.cos(phi), s * np.sin(phi), z)


@dataclass
class Problem:
    name: str
    nx: int
    ny: int
    nz: int
    rho0: np.ndarray          # slot order (x fastest)
    F: np.ndarray
    fixed: np.ndarray
    params: dict

    @property
    def n_dofs(self):
        return 3 * (self.nx + 1) * (self.ny + 1) * (self.nz + 1)

    @property
    def n_elements(self):
        return self.nx * self.ny * self.nz


def cantilever_problem(nx=30, ny=10, nz=10):
    """30x10x10 cantilever: x=0 face clamped, unit -z load at the far-end bottom-mid node;
    E=1, nu=0.3, Emin=1e-9, p=3, vf=0.3, rmin=1.5, move=0.2, all-design start at 0.3."""
    n_dofs = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    F = np.zeros(n_dofs)
    F[3 * node_id(nx, ny // 2, 0, ny, nz) + 2] = -1.0
    fixed = np.array([3 * node_id(0, j, k, ny, nz) + c
                      for j in range(ny + 1) for k in range(nz + 1) for c in range(3)],
                     dtype=np.int64)
    rho0 = np.full(nx * ny * nz, 0.3)
    params = dict(element_size=1.0, E=1.0, nu=0.3, E_min=1e-9, penalty=3.0,
                  volume_fraction=0.3, filter_radius=1.5, move_limit=0.2, rho_min=0.01)
    return Problem("cantilever_%dx%dx%d" % (nx, ny, nz), nx, ny, nz, rho0, F, fixed, params)


def bracket_like_problem(nx=120, ny=60, nz=60, seed=2, n_loads=3, volfrac=0.25, rmin=2.0):
    """Bracket-like domain: two solid 'bolt' pads on the x=0 face (clamped there), one void
    box of seeded size inside the design space, seeded point loads on the x=nx face with
    uniform unit direction and |F|~U(0.5,1). Sizes scale with nx/120."""
    sc = nx / 120.0
    rng = SplitMix64(seed)
    g = np.zeros((nx, ny, nz), dtype=np.int8)
    pad = max(1, int(round(10 * sc)))
    y_lo = [int(round(5 * sc)), int(round(45 * sc))]
    z_lo = int(round(25 * sc))
    for y0 in y_lo:
        g[0:pad, y0:y0 + pad, z_lo:z_lo + pad] = FIXED
    sx, sy, sz = (max(1, int(round(rng.randint(10, 20) * sc))) for _ in range(3))
    bx0, bx1 = int(round(30 * sc)), int(round(90 * sc))
    by0, by1 = int(round(15 * sc)), int(round(45 * sc))
    bz0, bz1 = int(round(15 * sc)), int(round(45 * sc))
    ox = rng.randint(bx0, max(bx0, bx1 - sx))
    oy = rng.randint(by0, max(by0, by1 - sy))
    oz = rng.randint(bz0, max(bz0, bz1 - sz))
    g[ox:ox + sx, oy:oy + sy, oz:oz + sz] = VOID
    n_dofs = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    F = np.zeros(n_dofs)
    for _ in range(n_loads):
        j, k = rng.randint(0, ny), rng.randint(0, nz)
        d = rng.unit_vector()
        mag = rng.uniform(0.5, 1.0)
        for c in range(3):
            F[3 * node_id(nx, j, k, ny, nz) + c] += mag * d[c]
    fixed = []
    for y0 in y_lo:
        for j in range(y0, y0 + pad + 1):
            for k in range(z_lo, z_lo + pad + 1):
                fixed.extend(3 * node_id(0, j, k, ny, nz) + c for c in range(3))
    dom = VoxelDomain(grid=g, resolution_mm=1.0)
    rho0 = dom.initial_density(volfrac, 0.0, order="slot")
    params = dict(element_size=1.0, E=1.0, nu=0.3, E_min=1e-9, penalty=3.0,
                  volume_fraction=volfrac, filter_radius=rmin, move_limit=0.2, rho_min=0.01)
    return Problem("bracket_like_%dx%dx%d_seed%d" % (nx, ny, nz, seed), nx, ny, nz, rho0, F,
                   np.unique(np.array(fixed, dtype=np.int64)), params)


def make_config(problem, **overrides):
    """SIMPConfig for a synthetic problem (reference SIMPConfig defaults otherwise)."""
    from .optimizer import SIMPConfig
    p = problem.params
    c = SIMPConfig(volume_fraction=p["volume_fraction"], penalty=p["penalty"],
                   filter_radius=p["filter_radius"], move_limit=p["move_limit"],
                   rho_min=p["rho_min"], E_min=p["E_min"], max_iterations=50,
                   convergence_tol=0.0, verbose=False, solver_tolerance=1e-8,
                   solver_max_iterations=100000, penalty_factor=1e8,
                   density_index="consistent")
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def _face_nodes(nx, ny, nz, face):
    """Node ids (z fastest) of one boundary face: 0/1 x=0/x=nx, 2/3 y=0/y=ny, 4/5 z=0/z=nz."""
    axis, side = face // 2, face % 2
    dims = (nx + 1, ny + 1, nz + 1)
    idx = [np.arange(d) for d in dims]
    idx[axis] = np.array([0 if side == 0 else dims[axis] - 1])
    I, J, K = np.meshgrid(*idx, indexing="ij")
    return (I * dims[1] * dims[2] + J * dims[2] + K).ravel()


def batch_problems(n=64, seed=3, lo=16, hi=64):
    """The dataset-generation set: n problems, grid size per axis U{lo..hi}, one random face
    clamped (all DOFs of its nodes), 1-4 point loads on random free nodes of the other faces
    (uniform unit direction, magnitude U(0.5, 1)), volume fraction U(0.2, 0.5), filter radius
    U(1.2, 2.5); E=1, nu=0.3, Emin=1e-9, p=3, move=0.2, rho_min=0.01; initial density = the
    volume fraction everywhere (all design). Drawn from one splitmix64 stream."""
    rng = SplitMix64(seed)
    out = []
    for t in range(n):
        nx, ny, nz = (rng.randint(lo, hi) for _ in range(3))
        n_dofs = 3 * (nx + 1) * (ny + 1) * (nz + 1)
        clamp = rng.randint(0, 5)
        clamped = _face_nodes(nx, ny, nz, clamp)
        fixed = np.sort(np.concatenate([3 * clamped + c for c in range(3)])).astype(np.int64)
        F = np.zeros(n_dofs)
        others = [f for f in range(6) if f != clamp]
        for _ in range(rng.randint(1, 4)):
            nodes = np.setdiff1d(_face_nodes(nx, ny, nz, others[rng.randint(0, 4)]), clamped)
            node = int(nodes[rng.randint(0, len(nodes) - 1)])
            d = rng.unit_vector()
            mag = rng.uniform(0.5, 1.0)
            for c in range(3):
                F[3 * node + c] += mag * d[c]
        vf = rng.uniform(0.2, 0.5)
        rmin = rng.uniform(1.2, 2.5)
        params = dict(volume_fraction=vf, penalty=3.0, filter_radius=rmin, move_limit=0.2,
                      rho_min=0.01, E_min=1e-9)
        out.append(Problem("batch%d_%dx%dx%d" % (t, nx, ny, nz), nx, ny, nz,
                           np.full(nx * ny * nz, vf), F, fixed, params))
    return out
