"""CPU tests: pin the oracle (the checker) against the reference's own artefacts, closed
forms, brute force and invariants. No GPU needed."""
import json
import os

import numpy as np
import pytest

import oracle_ffi as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_runs.json")))


# ---- Ke0 (element_kernel.cpp:155-179) ------------------------------------------------------
def test_ke0_closed_form_values():
    K = O.ke0(1.0, 0.3, 1.0)
    # exact rationals for E=1, nu=0.3, h=1 (2-point Gauss is exact for this integrand)
    assert K[0, 0] == pytest.approx(55 / 234, rel=1e-15)
    assert np.trace(K) == pytest.approx(220 / 39, rel=1e-15)
    assert K[0, 1] == pytest.approx(0.0801282051282051, rel=1e-13)
    assert K[0, 3] == pytest.approx(-0.1068376068376068, rel=1e-13)


def test_ke0_invariants():
    # main.cpp:50-64: symmetric, PSD; rigid-body null space of dimension 6
    K = O.ke0(113.8e9, 0.342, 0.004)
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
    ev = np.linalg.eigvalsh(K)
    assert ev[0] > -1e-6 * ev[-1]
    assert np.sum(np.abs(ev) < 1e-10 * ev[-1]) == 6
    assert ev[6] > 1e-3 * ev[-1]
    # rigid translations and rotations are in the null space
    x = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    for c in range(3):
        t = np.zeros(24)
        t[c::3] = 1.0
        assert np.abs(K @ t).max() < 1e-9 * np.abs(K).max()
    rot = np.zeros(24)
    rot[0::3], rot[1::3] = -x[:, 1], x[:, 0]
    assert np.abs(K @ rot).max() < 1e-9 * np.abs(K).max()


def test_ke0_scaling_linear_in_E_h():
    a = O.ke0(1.0, 0.3, 1.0)
    b = O.ke0(7.0, 0.3, 0.25)
    assert np.allclose(b, 7.0 * 0.25 * a, rtol=1e-14, atol=1e-16)


def test_ke0_lambda_max():
    ev = np.linalg.eigvalsh(O.ke0(1.0, 0.3, 1.0))
    assert ev[-1] == pytest.approx(1.25, rel=1e-13)


# ---- DOF map (element_kernel.cpp:186-216) ---------------------------------------------------
def test_element_dofs_known_case():
    # the case printed by main.cpp:204-211
    assert O.element_dofs(0, 0, 0, 10, 10, 10).tolist() == [
        0, 1, 2, 363, 364, 365, 396, 397, 398, 33, 34, 35,
        3, 4, 5, 366, 367, 368, 399, 400, 401, 36, 37, 38]


# ---- filter (topology_optimizer.cpp:60-119) -------------------------------------------------
@pytest.mark.parametrize("r,S", [(1.2, 7), (1.5, 19), (2.0, 33), (2.5, 81), (3.5, 179)])
def test_filter_interior_stencil_size(r, S):
    offs, idx, w = O.build_filter(11, 11, 11, r)
    e = 5 * 121 + 5 * 11 + 5
    assert offs[e + 1] - offs[e] == S


def test_filter_average_neighbors_notebook():
    # notebook cell 23 output "Average neighbors: 14.6" for 10x5x6, r=1.5
    offs, _, _ = O.build_filter(10, 5, 6, 1.5)
    avg = offs[-1] / 300
    assert avg == pytest.approx(14.613333333333333, rel=1e-15)
    assert round(avg, 1) == GOLD["notebook_cell23"]["average_neighbors"]


def test_filter_lists_brute_force():
    nx, ny, nz, r = 5, 4, 3, 1.8
    offs, idx, w = O.build_filter(nx, ny, nz, r)
    pts = np.array([(x, y, z) for z in range(nz) for y in range(ny) for x in range(nx)])
    for e in range(nx * ny * nz):
        d = np.sqrt(((pts - pts[e]) ** 2).sum(1))
        nb = np.nonzero(d <= r)[0]  # x-fastest order == dz, dy, dx loop order
        assert idx[offs[e]:offs[e + 1]].tolist() == nb.tolist()
        assert np.array_equal(w[offs[e]:offs[e + 1]], np.maximum(0.0, r - d[nb]))


def test_filter_constant_field_is_identity():
    nx, ny, nz = 6, 5, 4
    offs, idx, w = O.build_filter(nx, ny, nz, 2.0)
    rho = np.full(nx * ny * nz, 0.4)
    dc = np.full(nx * ny * nz, -2.5)
    assert np.allclose(O.apply_filter(offs, idx, w, dc, rho), -2.5, rtol=1e-15)


def test_filter_fallback_zero_density():
    offs, idx, w = O.build_filter(4, 4, 4, 1.5)
    dc = np.linspace(-1, -2, 64)
    out = O.apply_filter(offs, idx, w, dc, np.zeros(64))
    assert np.array_equal(out, dc)


# ---- bracket set-up (am_topopt.cpp:98-334) --------------------------------------------------
def test_bracket_counts_cpp_test_run():
    g, (nx, ny, nz) = O.bracket(4.0)
    G = GOLD["cpp_test_run"]
    assert [nx, ny, nz] == G["domain_shape"]
    assert ((g == 0).sum(), (g == 1).sum(), (g == -1).sum()) == (
        G["n_design"], G["n_fixed"], G["n_void"])


def test_bracket_notebook_case():
    g, (nx, ny, nz) = O.bracket(12.0)
    G = GOLD["notebook_cell23"]
    assert [nx, ny, nz] == G["domain_shape"]
    assert 3 * (nx + 1) * (ny + 1) * (nz + 1) == G["n_dofs"]
    assert len(O.bracket_fixed(nx, ny, nz)) == G["n_fixed_dofs"]
    F, _ = O.bracket_forces(nx, ny, nz, 12.0)
    assert np.abs(F).max() == G["force_max"]
    rho = np.where(g == 0, 0.3, np.where(g == 1, 1.0, 0.0))
    assert round(rho.mean(), 4) == G["initial_density_mean"]


# ---- assembly, operator, BCs ---------------------------------------------------------------
def _random_rho(n, seed, lo=0.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, n)


@pytest.mark.parametrize("ref_compat", [0, 1])
def test_csr_equals_matrix_free(ref_compat):
    nx, ny, nz = 5, 4, 3
    rho = _random_rho(nx * ny * nz, 1)
    rp, col, val = O.assemble_csr(nx, ny, nz, rho, ref_compat=ref_compat)
    K = O.ke0()
    s = np.where(rho < 0.01, 0.0, 1e-9 + rho ** 3 * (1 - 1e-9))
    if not ref_compat:  # element (i,j,k) reads slot k*nx*ny + j*nx + i
        s = s.reshape(nz, ny, nx).transpose(2, 1, 0).ravel()
    x = np.random.default_rng(2).standard_normal(len(rp) - 1)
    y1 = O.spmv(rp, col, val, x)
    y2 = O.matfree(nx, ny, nz, K, s, x)
    assert np.abs(y1 - y2).max() <= 1e-13 * np.abs(y1).max()


def test_csr_symmetric_and_rigid_null():
    nx, ny, nz = 4, 3, 3
    rp, col, val = O.assemble_csr(nx, ny, nz, _random_rho(36, 3, 0.1, 1.0))
    n = len(rp) - 1
    A = np.zeros((n, n))
    for r in range(n):
        A[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()   # main.cpp:282-298
    for c in range(3):
        t = np.zeros(n)
        t[c::3] = 1.0
        assert np.abs(A @ t).max() <= 1e-12 * np.abs(A).max()


def test_penalty_closed_form():
    nx, ny, nz = 3, 2, 2
    rp, col, val = O.assemble_csr(nx, ny, nz, np.full(12, 0.5))
    d0 = np.array([val[rp[r]:rp[r + 1]][col[rp[r]:rp[r + 1]] == r][0] for r in range(len(rp) - 1)])
    F = np.ones(len(rp) - 1)
    fixed = np.array([0, 1, 2, 5, 5], dtype=np.int64)  # DOF 5 listed twice
    P = O.apply_penalty(rp, col, val, F, fixed, 1e8)
    assert P == 1e8 * np.abs(d0).max()
    d1 = np.array([val[rp[r]:rp[r + 1]][col[rp[r]:rp[r + 1]] == r][0] for r in range(len(rp) - 1)])
    assert d1[0] == d0[0] + P and d1[5] == (d0[5] + P) + P and d1[3] == d0[3]
    assert F[[0, 1, 2, 5]].tolist() == [0, 0, 0, 0] and F[3] == 1.0
    with pytest.raises(ValueError):
        O.run(O.make_config(2, 2, 2, max_iterations=1), np.full(8, 0.3),
              np.zeros(81), np.array([81]))


def test_pcg_matches_dense_solve():
    nx, ny, nz = 4, 2, 2
    rho = _random_rho(16, 4, 0.2, 1.0)
    rp, col, val = O.assemble_csr(nx, ny, nz, rho)
    n = len(rp) - 1
    F = np.random.default_rng(5).standard_normal(n)
    fixed = np.arange(3 * 9, dtype=np.int64)  # x=0 face: nodes 0..8
    O.apply_penalty(rp, col, val, F, fixed, 1e8)
    u, st = O.pcg(rp, col, val, F, tol=1e-12, maxit=5000)
    A = np.zeros((n, n))
    for r in range(n):
        A[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    ud = np.linalg.solve(A, F)
    assert st["converged"] and st["true_rel"] < 1e-11
    assert np.abs(u - ud).max() <= 1e-9 * np.abs(ud).max()


def test_pcg_zero_rhs_and_maxit():
    rp, col, val = O.assemble_csr(2, 2, 2, np.full(8, 0.5))
    u, st = O.pcg(rp, col, val, np.zeros(len(rp) - 1))
    assert st["converged"] and st["iterations"] == 0 and not u.any()
    F = np.random.default_rng(0).standard_normal(len(rp) - 1)
    O.apply_penalty(rp, col, val, F, np.arange(27, dtype=np.int64))
    _, st = O.pcg(rp, col, val, F, tol=1e-12, maxit=1)
    assert not st["converged"] and st["iterations"] == 1


# ---- sensitivities vs finite differences ----------------------------------------------------
def test_sensitivity_matches_finite_difference():
    # dc = -p rho^(p-1) u^T Ke0 u is dc/drho of c = F.u for s = rho^p (Emin = 0 exactly);
    # the reference omits the (1-Emin) factor, which is exact here.
    nx, ny, nz = 3, 2, 2
    rho = _random_rho(12, 7, 0.3, 0.9)
    fixed = np.arange(3 * 9, dtype=np.int64)
    n = 3 * 4 * 3 * 3
    F = np.zeros(n)
    F[-1] = -1.0
    F[-5] = 0.5

    def compliance(r):
        rp, col, val = O.assemble_csr(nx, ny, nz, r, Emin=0.0, rho_min=0.0, ref_compat=0)
        Fb = F.copy()
        O.apply_penalty(rp, col, val, Fb, fixed, 1e8)
        u, _ = O.pcg(rp, col, val, Fb, tol=1e-14, maxit=10000)
        return u @ F, u

    c0, u = compliance(rho)
    dc = O.sensitivities(nx, ny, nz, O.ke0(), u, rho)
    for e in (0, 5, 11):
        h = 1e-6
        rp_, rm_ = rho.copy(), rho.copy()
        rp_[e] += h
        rm_[e] -= h
        fd = (compliance(rp_)[0] - compliance(rm_)[0]) / (2 * h)
        assert dc[e] == pytest.approx(fd, rel=2e-6)


# ---- OC (topology_optimizer.cpp:205-287) ----------------------------------------------------
def test_oc_invariants():
    n = 500
    rng = np.random.default_rng(9)
    rho = rng.uniform(0.05, 0.95, n)
    dcf = -rng.uniform(0.01, 3.0, n)
    design = (rng.uniform(size=n) > 0.2).astype(np.uint8)
    new, ch, npass, lam = O.oc_update(rho, dcf, design, 0.4, 0.2, 0.01)
    d = design.astype(bool)
    assert np.array_equal(new[~d], rho[~d])
    assert np.all(new[d] >= np.maximum(0.01, rho[d] - 0.2) - 1e-15)
    assert np.all(new[d] <= np.minimum(1.0, rho[d] + 0.2) + 1e-15)
    assert abs(new[d].mean() - 0.4) < 1e-3
    assert ch == np.abs(new - rho).max()
    assert 0 < npass <= 100

    # the lambda found is where vol(lambda) crosses vf: monotone, bracketed within 1e-4
    def vol(lm):
        b = np.sqrt(-dcf[d] / (lm + 1e-20))
        c = np.maximum(0.01, np.maximum(rho[d] - 0.2, np.minimum(1.0, np.minimum(rho[d] + 0.2, rho[d] * b))))
        return c.mean()
    assert (vol(lam * (1 - 2e-4)) - 0.4) * (vol(lam * (1 + 2e-4)) - 0.4) <= 0


def test_oc_positive_sensitivity_nan_semantics():
    # dc > 0 gives Be = NaN; std::min(rho+move, NaN) keeps rho+move (topology_optimizer.cpp:257)
    rho = np.array([0.5, 0.5, 0.5, 0.5])
    dcf = np.array([1.0, -1.0, -1.0, -1.0])
    new, _, _, _ = O.oc_update(rho, dcf, np.ones(4, np.uint8), 0.5, 0.2, 0.01)
    assert new[0] == 0.7


# ---- end-to-end against the reference's artefacts -------------------------------------------
def test_reproduces_cpp_test_run():
    """data/cpp_test_run (am_topopt --resolution 4 --iterations 5): compliance and volume
    histories to all printed digits, density field to 1e-6 (the reference's IC-CG tol 1e-6
    vs this Jacobi-PCG at 1e-10)."""
    G = GOLD["cpp_test_run"]
    g, (nx, ny, nz) = O.bracket(4.0)
    rho0 = np.where(g == 0, 0.25, np.where(g == 1, 1.0, 1e-4))
    F, _ = O.bracket_forces(nx, ny, nz, 4.0)
    fixed = O.bracket_fixed(nx, ny, nz)
    cfg = O.make_config(nx, ny, nz, h=0.004, E=113.8e9, nu=0.342, volume_fraction=0.25,
                        filter_radius=3.5, move_limit=0.1, rho_min=1e-4, E_min=1e-6,
                        max_iterations=5, convergence_tol=0.01, pcg_tol=1e-10,
                        pcg_maxit=50000)
    rho, hist, conv, _ = O.run(cfg, rho0, F, fixed)
    assert not conv and len(hist) == 5
    for h, c, v in zip(hist, G["compliance"], G["volume"]):
        assert float("%g" % h["compliance"]) == c
        assert round(h["volume_fraction"], 6) == v
    ref = np.load(os.path.join(HERE, "golden", "cpp_test_run_density_field.npy"),
                  allow_pickle=False)
    mine = rho.reshape(nz, ny, nx).transpose(2, 1, 0)  # saveNpy transposition (am_topopt.cpp:378-390)
    assert np.abs(mine - ref).max() < 1e-6


def test_notebook_iteration0_compliance():
    G = GOLD["notebook_cell23"]
    g, (nx, ny, nz) = O.bracket(12.0)
    gC = g.reshape(nz, ny, nx).transpose(2, 1, 0).ravel()   # density passed in C order
    rho0 = np.where(gC == 0, 0.3, np.where(gC == 1, 1.0, 0.0))
    F, _ = O.bracket_forces(nx, ny, nz, 12.0)
    fixed = O.bracket_fixed(nx, ny, nz)
    cfg = O.make_config(nx, ny, nz, h=0.012, E=113.8e9, nu=0.342, volume_fraction=0.3,
                        filter_radius=1.5, move_limit=0.1, rho_min=0.01, E_min=1e-3,
                        max_iterations=1, pcg_tol=1e-12, pcg_maxit=50000)
    _, hist, _, _ = O.run(cfg, rho0, F, fixed)
    assert hist[0]["compliance"] == pytest.approx(G["compliance_iter0"], rel=1e-9)


def test_oc_exact_tie_goes_to_l2():
    """The bisection's first pass lands exactly on vf (oracle_ffi.oc_tie_case): the strict
    comparison sends the tie to l2, so the final lambda lies below the first midpoint and the
    element ends above 0.75; the move limit (0.375 + 0.5) and the bisection tolerance bound it."""
    rho, dcf, design, vf, move, rho_min = O.oc_tie_case()
    new, ch, npass, lam = O.oc_update(rho, dcf, design, vf, move, rho_min)
    lam1 = 0.5 * (1e-20 + 1e20)
    assert lam < lam1
    assert 0.75 < new[5] < 0.75 * (1 + 1e-4)
    assert np.array_equal(np.delete(new, 5), np.delete(rho, 5))  # non-design unchanged
    assert 10 <= npass <= 20
