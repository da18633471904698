"""GPU tests of the matrix-core operator (element-batched Ke0 x_e on v_mfma_f32_16x16x16_bf16
with x and Ke0 split into two BF16 terms each) and of the solvers that use it."""
import numpy as np
import pytest

import am_ml_suite_amd as A
from am_ml_suite_amd.problems import make_config
import oracle_ffi as O

pytestmark = pytest.mark.gpu

MAT = A.MaterialProperties(1.0, 0.3)


def _opt(p, **kw):
    opt = A.TopologyOptimizer(p.nx, p.ny, p.nz, 1.0, MAT, make_config(p, **kw))
    opt.set_loads(p.F, p.fixed)
    opt.set_density(p.rho0)
    return opt


@pytest.mark.parametrize("shape", [(13, 7, 5), (40, 20, 20), (2, 1, 1)])
def test_mfma_operator_matches_fp64(shape):
    """Random and smooth fields, with and without penalty rows, on grids with partial tiles:
    the split-BF16 MFMA operator is within 3e-5 relative (L2) of the FP64 vector operator on
    random fields (3e-4 on a smooth one) and agrees with the oracle's assembled CSR SpMV."""
    p = A.bracket_like_problem(*shape)
    opt = _opt(p)
    n = opt.n_dofs
    rng = np.random.default_rng(1)
    i, j, k = np.meshgrid(np.arange(p.nx + 1), np.arange(p.ny + 1), np.arange(p.nz + 1),
                          indexing="ij")
    smooth = np.stack([np.sin(0.3 * i + 0.1 * j), np.cos(0.2 * k) * j, 0.01 * i * k],
                      axis=-1).ravel()
    # x and Ke0 are carried to ~2^-17 relative, so the error scales with |K||x|: a smooth
    # field, whose K x cancels more, gets a looser bound relative to |K x|
    for x, tol in ((rng.standard_normal(n), 3e-5), (smooth, 3e-4)):
        for bc in (False, True):
            y64 = opt.apply_operator(x, with_bc=bc)
            ymf = opt.apply_operator(x, with_bc=bc, precision="mixed_bf16")
            y32 = opt.apply_operator(x, with_bc=bc, precision="mixed_fp32")
            e_mf = np.linalg.norm(ymf - y64) / np.linalg.norm(y64)
            e_32 = np.linalg.norm(y32 - y64) / np.linalg.norm(y64)
            print("shape %s bc %d: mfma %.2e fp32 %.2e" % (shape, bc, e_mf, e_32))
            assert e_mf <= tol, (bc, e_mf, e_32)
    # element scales as the FP64 path sees them: compare with the oracle's CSR without BCs
    rp, col, val = O.assemble_csr(p.nx, p.ny, p.nz, _ref_order(p))
    x = rng.standard_normal(n)
    yref = O.spmv(rp, col, val, x)
    ymf = opt.apply_operator(x, with_bc=False, precision="mixed_bf16")
    assert np.linalg.norm(ymf - yref) <= 3e-5 * np.linalg.norm(yref)


def _ref_order(p):
    """The density in the oracle's C-order (the consistent index reads slot order)."""
    return p.rho0.reshape(p.nz, p.ny, p.nx).transpose(2, 1, 0).ravel()


def test_mfma_multigrid_full_size():
    """1,350,723-DOF bracket: flexible CG with the V-cycle's fine-level smoothing on the matrix
    cores reaches the 1e-8 FP64 true residual with compliance within 1e-6 of plain FP64
    Jacobi-PCG, in as few iterations as the FP32-smoothed V-cycle (+10%)."""
    p = A.bracket_like_problem()
    out = {}
    for name, kw in (("jacobi", dict()), ("mg32", dict(preconditioner="multigrid")),
                     ("mgbf16", dict(preconditioner="multigrid", precision="mixed_bf16"))):
        opt = _opt(p, **kw)
        u, st = opt.solve()
        assert st["converged"] and st["true_rel_residual"] <= 1.5e-8, (name, st)
        out[name] = (u @ p.F, st["iterations"])
        opt.close()
    assert out["mgbf16"][0] == pytest.approx(out["jacobi"][0], rel=1e-6)
    assert out["mgbf16"][1] <= 1.1 * out["mg32"][1] + 2
    print("iterations: jacobi %d, multigrid fp32 %d, multigrid mfma %d"
          % (out["jacobi"][1], out["mg32"][1], out["mgbf16"][1]))


def test_mfma_refinement_solver():
    """Jacobi-PCG inner solves on the MFMA operator inside FP64 residual refinement: the FP64
    true residual meets 1e-8 and compliance matches plain FP64 within 1e-6."""
    p = A.bracket_like_problem(40, 20, 20)
    u64, s64 = _opt(p).solve()
    u, st = _opt(p, precision="mixed_bf16").solve()
    assert st["converged"] and st["true_rel_residual"] <= 1.5e-8, st
    assert (u @ p.F) == pytest.approx(u64 @ p.F, rel=1e-6)


def test_mfma_multigrid_run_matches_oracle():
    """OC loop with the MFMA-smoothed multigrid solver on the 30x10x10 cantilever: compliance
    within 1e-6 of the oracle's Jacobi-PCG path for 20 iterations, density within 1e-4."""
    p = A.cantilever_problem()
    its = 20
    opt = A.TopologyOptimizer(p.nx, p.ny, p.nz, 1.0, MAT,
                              make_config(p, preconditioner="multigrid", precision="mixed_bf16",
                                          max_iterations=its))
    res = opt.run(p.rho0, p.F, p.fixed)
    ocfg = O.make_config(p.nx, p.ny, p.nz, max_iterations=its, pcg_tol=1e-8, pcg_maxit=100000,
                         ref_compat=0)
    rho_ref, hist, _, _ = O.run(ocfg, p.rho0, p.F, p.fixed)
    c = np.array(res.compliance_history)
    cr = np.array([h["compliance"] for h in hist])
    assert np.abs(c / cr - 1).max() <= 1e-6
    assert np.abs(res.final_density - rho_ref).max() <= 1e-4
