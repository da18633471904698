"""GPU parity of the Python SIMP flavour (am_ml_suite_amd.simp, amt_simp_*) against the
numpy/scipy oracle (oracle/simp_py.py) on a grid with fixed, void and design cells, orphan
DOFs and two load cases with different constrained faces."""
import os
import sys

import numpy as np
import pytest

from am_ml_suite_amd.problems import VoxelDomain
from am_ml_suite_amd.simp import MaterialProperties, SIMPOptimizer, SIMPParams
from simp_py_cases import LoadCase, small_problem

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import simp_py as S  # noqa: E402

pytestmark = pytest.mark.gpu

RES_MM = 2.0


def _pair(vf=0.3, r=1.5, tol=0.01, max_it=10, solve_tol=1e-12, shape=(12, 6, 8)):
    grid, loads = small_problem(*shape)
    prm = SIMPParams(volume_fraction=vf, filter_radius=r, convergence_tol=tol,
                     max_iterations=max_it, solve_tol=solve_tol)
    opt = SIMPOptimizer(VoxelDomain(grid, RES_MM), [LoadCase(F, c) for F, c in loads],
                        MaterialProperties(), prm, verbose=False)
    P = S.Problem(grid, RES_MM, loads, volume_fraction=vf, filter_radius=r,
                  convergence_tol=tol)
    return opt, P


@pytest.mark.parametrize("r", [1.2, 2.0, 3.5])
def test_filter_matches_oracle(r):
    """convolve(x, cone, mode='reflect') on the device vs scipy, radius up to 3.5 (the box
    reaches past the grid's thin axes, so the extension reflects more than once)."""
    opt, P = _pair(r=r, shape=(9, 3, 5))
    x = np.random.default_rng(2).standard_normal(opt.density.shape)
    y = opt.apply_filter(x)
    ref = S.apply_filter(x, P.w)
    assert np.abs(y - ref).max() <= 1e-14 * np.abs(ref).max()
    opt.close()


def test_one_iteration_matches_oracle():
    """One loop pass from the uniform start: per-load-case displacements (orphan and
    constrained DOFs exactly 0), compliance, volume and the OC density."""
    opt, P = _pair()
    d0 = opt.density.copy()
    new, it = opt.step(d0)
    ref, info = S.step(P, d0)
    print("compliance gpu %.15e oracle %.15e, true residual %.2e, pcg %d"
          % (it.compliance, info["compliance"], it.max_true_rel_residual, it.solver_iterations))
    assert it.solver_converged and it.max_true_rel_residual <= 1e-11
    assert it.compliance == pytest.approx(info["compliance"], rel=1e-9)
    nx, ny, nz = P.grid.shape
    corner = 3 * ((nx * (ny + 1) + ny) * (nz + 1) + nz)  # node touched by void cells only
    for lc, u_ref in enumerate(info["u"]):
        u = opt.displacement(lc)
        assert np.all(u[P.loads[lc][1]] == 0) and np.all(u[corner:corner + 3] == 0)
        assert np.linalg.norm(u - u_ref) <= 1e-9 * np.linalg.norm(u_ref)
    assert it.volume == pytest.approx(info["volume"], rel=1e-12)
    assert it.change == pytest.approx(info["change"], rel=1e-9)
    assert np.abs(new - ref).max() <= 1e-9
    opt.close()


def test_run_matches_oracle():
    """SIMPOptimizer.run for up to 12 iterations against the oracle's loop: compliance and
    volume history, stopping iteration and the final density."""
    opt, P = _pair(max_it=12, tol=0.02)
    res = opt.run()
    ref, hist = S.run(P, 12)
    c = np.array(res.compliance_history)
    cr = np.array([h["compliance"] for h in hist])
    print("iterations gpu %d oracle %d; compliance %s" % (res.iterations, len(hist), c))
    assert res.iterations == len(hist)
    assert res.converged == (hist[-1]["change"] < P.tol)
    assert np.abs(c / cr - 1).max() <= 1e-8
    assert np.abs(np.array(res.volume_history) - [h["volume"] for h in hist]).max() <= 1e-10
    assert np.abs(res.density - ref).max() <= 1e-7
    assert max(res.true_rel_residual) <= 1e-11
    opt.close()


def test_design_volume_and_bounds():
    """After every step the design elements stay in [rho_min, 1] and within the move limit,
    fixed/void elements keep their value, and the design volume sits at the target within
    the bisection tolerance."""
    opt, P = _pair(vf=0.25, r=2.0)
    d = opt.density.copy()
    grid = P.grid
    for _ in range(4):
        new, it = opt.step(d)
        des = grid == 0
        assert np.all(new[des] >= 0.01 - 1e-15) and np.all(new <= 1.0)
        assert np.all(np.abs(new - d) <= 0.2 + 1e-12)
        assert np.array_equal(new[~des], d[~des])
        assert abs(it.volume - 0.25) < 5e-3
        d = new
    opt.close()
