"""Small problems for the Python SIMP flavour: a voxel grid with fixed, void and design cells
(including a void corner whose nodes have no stiffness: orphan DOFs) and two load cases with
different constrained faces."""
import numpy as np


def small_problem(nx=12, ny=6, nz=8):
    grid = np.zeros((nx, ny, nz), dtype=np.int8)
    grid[:, :, 0] = 1                                   # fixed base layer
    grid[nx // 3:nx // 3 + 4, ny // 3:ny // 3 + 2, 3:6] = -1  # interior void box
    grid[nx - 2:, ny - 2:, nz - 2:] = -1                # void corner: orphan nodes
    nn = (nx + 1, ny + 1, nz + 1)
    n_dof = 3 * nn[0] * nn[1] * nn[2]

    def node(i, j, k):
        return (i * nn[1] + j) * nn[2] + k

    # case 1: base (k = 0) clamped, -z loads on top nodes
    F1 = np.zeros(n_dof)
    for i, j in ((nx // 2, ny // 2), (nx // 4, ny // 2), (nx - 3, 1)):
        F1[3 * node(i, j, nz) + 2] -= 1000.0
    c1 = [3 * node(i, j, 0) + c for i in range(nn[0]) for j in range(nn[1]) for c in range(3)]
    # case 2: x = 0 face clamped, +y / -x loads on the far face
    F2 = np.zeros(n_dof)
    F2[3 * node(nx, ny // 2, nz // 2) + 1] += 700.0
    F2[3 * node(nx, 0, 1) + 0] -= 400.0
    c2 = [3 * node(0, j, k) + c for j in range(nn[1]) for k in range(nn[2]) for c in range(3)]
    return grid, [(F1, np.array(c1, dtype=np.int64)), (F2, np.array(c2, dtype=np.int64))]


class LoadCase:
    """get_force_vector / get_constrained_dofs, as SIMPOptimizer consumes them."""

    def __init__(self, F, cons):
        self.F, self.cons = F, cons

    def get_force_vector(self, n_dofs):
        assert n_dofs == self.F.size
        return self.F.copy()

    def get_constrained_dofs(self):
        return np.unique(self.cons)
