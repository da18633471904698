"""CPU checks that pin the Python-SIMP oracle (oracle/simp_py.py) before it judges the GPU."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import simp_py as S  # noqa: E402
from simp_py_cases import small_problem  # noqa: E402


def _reflect(m, n):
    p = 2 * n
    m %= p
    return m if m < n else p - 1 - m


@pytest.mark.parametrize("r,count", [(1.5, 19), (2.0, 27), (1.2, 7)])
def test_cone_kernel(r, count):
    """Weights r - dist inside the ball (zero on its surface), normalised to 1, with all the
    cube's symmetries; 19 / 27 / 7 lattice points strictly inside r = 1.5 / 2 / 1.2."""
    w = S.cone_kernel(r)
    assert w.sum() == pytest.approx(1.0, abs=1e-15)
    assert np.count_nonzero(w) == count
    for t in ((1, 0, 2), (2, 1, 0)):
        assert np.array_equal(w, w.transpose(t))
    assert np.array_equal(w, w[::-1, :, :])


@pytest.mark.parametrize("shape,r", [((5, 4, 3), 1.5), ((7, 6, 5), 2.0), ((3, 2, 4), 3.5),
                                     ((1, 5, 2), 2.5)])
def test_filter_is_reflect_convolution(shape, r):
    """scipy's mode='reflect' is the half-sample symmetric extension, periodic with 2n beyond
    one reflection (radius larger than the grid): brute force written out."""
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape)
    w = S.cone_kernel(r)
    R = w.shape[0] // 2
    y = np.zeros(shape)
    nx, ny, nz = shape
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                acc = 0.0
                for a in range(2 * R + 1):
                    for b in range(2 * R + 1):
                        for c in range(2 * R + 1):
                            acc += w[a, b, c] * x[_reflect(i + R - a, nx), _reflect(j + R - b, ny),
                                                  _reflect(k + R - c, nz)]
                y[i, j, k] = acc
    assert np.allclose(S.apply_filter(x, w), y, rtol=0, atol=1e-14)


def test_element_stiffness():
    """E=1, nu=0.3, h=1: Ke[0,0] = 55/234 (exact integration), symmetric, six rigid-body
    zero modes; Ke is linear in E*h."""
    K = S.element_stiffness(1.0, 1.0, 0.3)
    assert K[0, 0] == pytest.approx(55 / 234, rel=1e-14)
    assert np.allclose(K, K.T, atol=1e-15)
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-12) == 6 and ev.min() > -1e-12
    assert np.allclose(S.element_stiffness(2.0, 3.0, 0.3), 6.0 * K, atol=1e-13)


def test_solve_orphans_and_galerkin():
    """Orphan DOFs (only void elements around) are eliminated with the constrained ones; the
    direct solve satisfies K_ff u_f = F_f; c = sum_e E_ratio u_e^T Ke u_e equals F.u."""
    grid, loads = small_problem()
    P = S.Problem(grid, 2.0, loads, volume_fraction=0.3, filter_radius=1.5)
    rhof = np.clip(S.apply_filter(np.full(grid.shape, 0.3), P.w), P.rho_min, 1.0)
    K = S.assemble(P.grid, rhof, P.Ke, P.E, P.E_min, P.p)
    orphans = np.where(np.abs(K.diagonal()) < 1e-15)[0]
    nx, ny, nz = grid.shape
    corner = 3 * ((nx * (ny + 1) + ny) * (nz + 1) + nz)
    assert corner in orphans and len(orphans) >= 3
    for F, cons in loads:
        u, free = S.solve(K, F, cons)
        assert np.all(u[cons] == 0) and np.all(u[orphans] == 0)
        r = K[free, :][:, free] @ u[free] - F[free]
        assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(F[free])
        c, _ = S.compliance(u, P.grid, rhof, P.Ke, P.E, P.E_min, P.p)
        assert c == pytest.approx(F @ u, rel=1e-9)


def test_sensitivities_finite_difference():
    """dc_e = -p rho_e^(p-1) u_e^T Ke u_e against central differences of c(rho) (no filter;
    the reference omits the (1 - E_min/E) factor, 1e-20 here)."""
    nx, ny, nz = 3, 2, 2
    grid = np.zeros((nx, ny, nz), dtype=np.int8)
    n_dof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    F = np.zeros(n_dof)
    F[3 * ((nx * (ny + 1) + 1) * (nz + 1) + 1) + 2] = -1.0
    cons = np.arange(3 * (ny + 1) * (nz + 1))
    Ke = S.element_stiffness(1.0, 1.0, 0.3)
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.3, 0.9, (nx, ny, nz))

    def c_of(r):
        K = S.assemble(grid, r, Ke, 1.0, 1e-9, 3.0)
        u, _ = S.solve(K, F, cons)
        return S.compliance(u, grid, r, Ke, 1.0, 1e-9, 3.0)

    c0, ce = c_of(rho)
    dc = S.sensitivities(ce, grid, rho, 3.0)
    for e in [(0, 0, 0), (1, 1, 0), (2, 0, 1)]:
        h = 1e-6
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h
        rm[e] -= h
        fd = (c_of(rp)[0] - c_of(rm)[0]) / (2 * h)
        assert dc[e] == pytest.approx(fd, rel=1e-6)


def test_oc_update_invariants():
    """Bounds, move limit, non-design elements untouched, and the final volume within the
    bisection's last bracket of the target."""
    rng = np.random.default_rng(5)
    shape = (6, 5, 4)
    rho = rng.uniform(0.05, 0.9, shape)
    dc = -rng.uniform(0.0, 5.0, shape)
    design = rng.uniform(size=shape) > 0.2
    new, passes = S.oc_update(rho, dc, design, 0.4, 0.2, 0.01)
    assert 10 < passes <= 100
    assert np.all(new[design] >= 0.01 - 1e-15) and np.all(new <= 1.0)
    assert np.all(np.abs(new - rho) <= 0.2 + 1e-12)
    assert np.array_equal(new[~design], rho[~design])
    assert abs(new[design].mean() - 0.4) < 2e-2
