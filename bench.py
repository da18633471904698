"""OC iterations/s on the 120x60x60 bracket-like problem (1,350,723 DOFs), every solve to a
1e-8 relative FP64 true residual, on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]

A step is one OC iteration (operator coefficients, PCG solve, compliance, sensitivities,
filter, OC bisection) of the SIMP loop, with all inputs resident in HBM. Two solver paths run
the same W + K OC iterations:
  - plain:  FP64 Jacobi-PCG from u = 0 (the reference's solve, Jacobi in place of IC);
  - fast:   FP64 flexible CG preconditioned by an FP32 Galerkin geometric-multigrid V-cycle,
            warm-started from the previous OC iteration's displacement, stopped on the same
            FP64 residual (the true residual is recomputed and reported).
`value` is the fast path's OC iterations/s; the plain path and the speed-up are reported
beside it, with the fast path's compliance/density error against the plain FP64 path.
With --paths ...,large the "large" object reports the 256x128x128 bracket (12,830,211 DOFs):
100 OC iterations with the V-cycle's fine-level smoothing on the matrix cores (split-BF16
MFMA operator, FP32 accumulation) against the same run with FP32 vector smoothing (FP64
flexible CG either way), with the density max-abs between them.
The "learned" object (default) reports the U-Net: warm_start trains it on the warm-start
residuals of 8 other 120x60x60 problems and times 12 OC iterations of the headline problem with
and without the learned step u += gamma M(r) before each multigrid solve; fcg_96x48x48 trains it
as a flexible-CG preconditioner on Jacobi-PCG residuals and solves a held-out 96x48x48 problem
against plain Jacobi-PCG and the multigrid FCG.
The "batch" object reports the dataset-generation workload: 64 generated problems of sizes
U{16..64}^3 (13.0M DOFs together) advanced in lockstep by the batched engine (FP32 matvec,
FP64 vectors and reductions, residual replacement, every solve to the 1e-8 true residual).
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per launch of the fast path's level-0 pre-smoothing with the flexible-CG update, by
# candidate, from the same kind of counter runs (profiles/fcg_apply_bracket120x60x60_hbm_traffic.md)
PRE_UPD_TRAFFIC_BYTES = {0: 101.34e6, 1: 107.04e6, 2: 89.50e6, 3: 84.41e6, 4: 100.79e6, 5: 87.98e6,
                         6: 84.38e6, 7: 100.07e6, 8: 87.94e6, 9: 84.39e6, 10: 86.62e6, 11: 93.69e6,
                         12: 88.60e6, 13: 102.36e6, 14: 86.58e6, 15: 95.60e6}
# HBM bytes per launch of the fast path's FP64 operator on this workload, by operator candidate
# (amt_config.operator_tile), from rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by a plain
# 256 MiB copy (profiles/fcg_apply_bracket120x60x60_hbm_traffic.md)
FCG_APPLY_TRAFFIC_BYTES = {0: 49.39e6, 1: 51.40e6, 2: 51.35e6, 3: 49.04e6, 4: 51.60e6, 5: 51.52e6,
                          6: 49.04e6, 7: 51.15e6, 8: 51.43e6, 9: 49.04e6, 10: 49.28e6, 11: 49.46e6,
                          12: 49.14e6, 13: 50.51e6, 14: 49.14e6, 15: 49.68e6}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_path(problem, warmup, steps, precision, warm_start, profile, precond="jacobi",
             operator_tile=-1):
    import am_ml_suite_amd as A
    from am_ml_suite_amd.problems import make_config
    p = problem
    cfg = make_config(p, precision=precision, warm_start=warm_start,
                      max_iterations=warmup + steps, preconditioner=precond,
                      operator_tile=operator_tile)
    opt = A.TopologyOptimizer(p.nx, p.ny, p.nz, 1.0, A.MaterialProperties(1.0, 0.3), cfg)
    rho = p.rho0.copy()
    opt.set_loads(p.F, p.fixed)
    opt.set_density(rho)
    opt.begin_run()
    res = []
    for it in range(warmup):
        r = opt.step()
        res.append(r)
        log("  [%s] warmup %d: c=%.9e pcg=%d outer=%d solve=%.3fs" % (
            precision, it, r.compliance, r.solver_iterations, r.outer_iterations, r.solve_time_s))
    if profile:
        opt.profile_enable(True)
    opt.sync()
    t0 = time.perf_counter()
    for it in range(steps):
        r = opt.step()
        res.append(r)
        log("  [%s] step %d: c=%.9e pcg=%d outer=%d res=%.2e solve=%.3fs wall=%.3fs" % (
            precision, it, r.compliance, r.solver_iterations, r.outer_iterations,
            r.true_rel_residual, r.solve_time_s, time.perf_counter() - t0))
    opt.sync()
    dt = time.perf_counter() - t0
    prof = opt.profile_read() if profile else (0, 0.0)
    prof_upd = opt.profile_read(1) if profile and precond == "multigrid" else (0, 0.0)
    dens = opt.get_density()
    op = opt.operator_info(0) if precond == "multigrid" else None
    op_upd = opt.operator_info(1) if precond == "multigrid" else None
    opt.close()
    return dict(time=dt, results=res, density=dens, prof=prof, op=op, prof_upd=prof_upd,
                op_upd=op_upd)


def run_large(steps):
    import am_ml_suite_amd as A
    t = time.perf_counter()
    p = A.bracket_like_problem(256, 128, 128, seed=5)
    log("large problem generated in %.1fs" % (time.perf_counter() - t))
    out = {}
    for name, prec in (("mfma", "mixed_bf16"), ("fp32_smoother", "fp64")):
        log("large [%s]: %s, %d DOFs" % (name, p.name, p.n_dofs))
        out[name] = run_path(p, 1, steps, prec, 2, False, "multigrid")
    m, f = out["mfma"], out["fp32_smoother"]
    t_m = m["results"][1:]
    cm = np.array([r.compliance for r in m["results"]])
    cf = np.array([r.compliance for r in f["results"]])
    return dict(workload=p.name, n_dofs=p.n_dofs, steps=steps,
                oc_iter_per_s=steps / m["time"], dof_per_s=p.n_dofs * steps / m["time"],
                fp32_smoother_oc_iter_per_s=steps / f["time"],
                density_maxabs_vs_fp64_path=float(np.abs(m["density"] - f["density"]).max()),
                max_rel_compliance_diff=float(np.abs(cm / cf - 1).max()),
                mean_fcg_iters=float(np.mean([r.solver_iterations for r in t_m])),
                max_true_rel_residual=float(max(r.true_rel_residual for r in t_m)),
                solver="FP64 flexible CG + FP32 V-cycle, fine-level smoothing on MFMA "
                       "(split-BF16 operator, FP32 accumulation)")


def run_simp_py(steps, cpu):
    """The Python SIMPOptimizer flavour (topopt.py:84-411) on the BRK-A-01 bracket: 4 mm
    (30x15x20, the reference's test-run resolution) and 2 mm (60x30x40); SIMPParams defaults,
    Ti6Al4V, each load-case solve to a 1e-10 relative residual. The oracle (numpy/scipy, sparse
    direct solve, as the reference) is timed on the 4 mm grid."""
    import am_ml_suite_amd as A
    from am_ml_suite_amd.simp import SIMPOptimizer, SIMPParams
    out = {}
    for res_mm in (4.0, 2.0):
        dom = A.bracket_domain(res_mm)
        lc = A.bracket_load_case(dom)
        opt = SIMPOptimizer(dom, [lc], params=SIMPParams(max_iterations=steps + 1),
                            verbose=False)
        d, it = opt.step(opt.density)  # warm-up
        first_c = it.compliance
        t0 = time.perf_counter()
        comp, pcg, resid = [], [], []
        for _ in range(steps):
            d, it = opt.step(d)
            comp.append(it.compliance)
            pcg.append(it.solver_iterations)
            resid.append(it.max_true_rel_residual)
        dt = time.perf_counter() - t0
        n_dof = 3 * (dom.shape[0] + 1) * (dom.shape[1] + 1) * (dom.shape[2] + 1)
        out["bracket_%gmm" % res_mm] = dict(
            grid=list(dom.shape), n_dofs=n_dof, steps=steps, oc_iter_per_s=steps / dt,
            mean_pcg_iters=float(np.mean(pcg)), max_true_rel_residual=float(max(resid)),
            first_compliance=first_c, last_compliance=comp[-1])
        opt.close()
        log("  [simp_py] %g mm: %.2f OC iter/s, %.0f PCG iterations/solve"
            % (res_mm, steps / dt, np.mean(pcg)))
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import simp_py as S
        dom = A.bracket_domain(4.0)
        lc = A.bracket_load_case(dom)
        P = S.Problem(dom.grid, 4.0, [(lc.get_force_vector(), lc.get_constrained_dofs())])
        d = np.full(dom.shape, P.vf)
        t0 = time.perf_counter()
        d, info = S.step(P, d)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = dict(value=1.0 / dt, unit="OC iter/s", cores=1, kind="port",
                                   sample="oracle/simp_py.py (numpy/scipy, SuperLU direct "
                                          "solve as the reference) on the 4 mm bracket, the "
                                          "first OC iteration from the uniform start",
                                   compliance=info["compliance"])
    return out


def run_ai(n_train=192, n_test=16, shape=(24, 16, 16), oc_iters=40, steps=1000, batch=8):
    """The AI density predictor: ground-truth designs from the GPU SIMP engine, U-Net training
    (device forward/backward, host Adam), held-out evaluation against the optimised designs."""
    from am_ml_suite_amd import ai
    probs = ai.ai_problems(n_train + n_test, shape=shape, seed=4)
    t0 = time.perf_counter()
    designs = ai.optimise(probs, oc_iters)
    t_gen = time.perf_counter() - t0
    log("  [ai] %d designs (%d OC iterations each) in %.1fs" % (len(probs), oc_iters, t_gen))
    samples = ai.make_samples(probs, designs)
    A = ai.AIOptimizer(shape=shape, seed=1)
    U = ai.AIOptimizer(shape=shape, seed=1)
    t0 = time.perf_counter()
    losses = A.train(samples[:n_train], steps, lr=1e-3, batch=batch, seed=0)
    t_train = time.perf_counter() - t0
    res_t, res_u, unif = [], [], []
    for p, d in zip(probs[n_train:], designs[n_train:]):
        res_t.append(A.validate_against_numerical(p, d))
        res_u.append(U.validate_against_numerical(p, d))
        unif.append(ai.compliance_of(p, np.full(d.shape, p.params["volume_fraction"]))
                    / res_t[-1]["compliance_numerical"])
    t0 = time.perf_counter()
    for p in probs[n_train:]:
        A.predict_density(ai._Domain(p), [ai._LC(p.F, p.fixed)], p.params["volume_fraction"])
    t_inf = (time.perf_counter() - t0) / n_test
    return dict(workload="ai_problems_%dx%dx%d_seed4" % shape, n_train=n_train, n_test=n_test,
                oc_iterations_per_design=oc_iters, numerical_s_per_design=t_gen / len(probs),
                train_steps=steps, batch=batch, train_s=t_train,
                loss_first=float(np.mean(losses[:10])), loss_last=float(np.mean(losses[-25:])),
                inference_ms_per_problem=1e3 * t_inf,
                iou_trained=float(np.mean([r["iou"] for r in res_t])),
                iou_untrained=float(np.mean([r["iou"] for r in res_u])),
                compliance_ratio_median_trained=float(np.median([r["compliance_ratio"] for r in res_t])),
                compliance_ratio_median_untrained=float(np.median([r["compliance_ratio"] for r in res_u])),
                compliance_ratio_median_uniform=float(np.median(unif)),
                compliance_ratios_trained=[round(r["compliance_ratio"], 3) for r in res_t])


def run_unet_conv(shape=(128, 64, 64), reps=20, rounds=3):
    """The U-Net inference forward on its own (14 brick convolutions, 3 poolings, the head) on
    the 128x64x64 grid, the 120x60x60 workload's node grid padded to multiples of 8: the best of
    `rounds` HIP-event timings of `reps` forwards, and its convolution FLOP rate."""
    import am_ml_suite_amd as A
    net = A.UNet(*shape, seed=3)
    n = shape[0] * shape[1] * shape[2]
    net.infer(np.random.default_rng(0).standard_normal((n, 16)).astype(np.float32) * 0.1)
    best_ms, best_tf = None, None
    for _ in range(rounds):
        ms, tf = net.benchmark(reps)
        if best_ms is None or ms < best_ms:
            best_ms, best_tf = ms, tf
    log("  [unet conv] %dx%dx%d forward %.4f ms, %.1f TFLOP/s (%.1f%% of dense BF16 peak)"
        % (shape + (best_ms, best_tf, 100 * best_tf / BF16_PEAK_TFLOPS)))
    return dict(grid="%dx%dx%d" % shape, forward_ms=best_ms, conv_tflops=best_tf,
                frac_bf16_peak=best_tf / BF16_PEAK_TFLOPS, forwards_timed=reps * rounds)


def run_learned(shape=(96, 48, 48), train_steps=400, samples=64):
    """The learned preconditioner on 96x48x48 bracket-like problems (698,691 DOFs): residual
    samples of a Jacobi-PCG solve on seed 4, GPU training, then on the held-out seed-5 problem
    the learned flexible CG (Jacobi-PCG fallback on stagnation) against plain Jacobi-PCG and
    the multigrid FCG, each to the 1e-8 FP64 residual."""
    import am_ml_suite_amd as A
    from am_ml_suite_amd.problems import make_config

    def make(seed, pre):
        p = A.bracket_like_problem(*shape, seed=seed)
        o = A.TopologyOptimizer(p.nx, p.ny, p.nz, 1.0, A.MaterialProperties(1.0, 0.3),
                                make_config(p, preconditioner=pre))
        o.set_loads(p.F, p.fixed)
        o.set_density(p.rho0)
        return o, p

    tr, _ = make(4, "unet")
    held, ph = make(5, "unet")
    net = A.UNet(optimizer=tr, seed=3)
    n = tr.collect_samples(samples)
    nh = held.collect_samples(samples)
    l0, j0 = held.train_unet(net, nh, lr=0.0, with_jacobi=True)
    t0 = time.perf_counter()
    losses, jl = tr.train_unet(net, train_steps, lr=1e-3, with_jacobi=True)
    t_train = time.perf_counter() - t0
    lh, jh = held.train_unet(net, nh, lr=0.0, with_jacobi=True)
    held.attach_unet(net)
    held.set_density(ph.rho0)
    held.sync()
    t0 = time.perf_counter()
    u, st = held.solve()
    t_l = time.perf_counter() - t0
    out = dict(workload="bracket_like_%dx%dx%d_seed5" % shape, n_dofs=ph.n_dofs,
               train_samples=n, train_steps=train_steps, train_s=t_train,
               heldout_energy_per_application_learned=float(np.mean(lh)),
               heldout_energy_per_application_jacobi=float(np.mean(jh)),
               learned=dict(iterations=st["iterations"], fallback=st["fallback"], time_s=t_l,
                            true_rel_residual=st["true_rel_residual"]))
    c_l = float(u @ ph.F)
    for name, pre in (("jacobi", "jacobi"), ("multigrid", "multigrid")):
        o, p = make(5, pre)
        o.sync()
        t0 = time.perf_counter()
        u2, s2 = o.solve()
        dt = time.perf_counter() - t0
        out[name] = dict(iterations=s2["iterations"], time_s=dt,
                         true_rel_residual=s2["true_rel_residual"],
                         compliance_rel_diff_learned=abs(c_l / float(u2 @ p.F) - 1))
        o.close()
    log("  [learned] held-out: learned %d its %.3fs (fallback %s), jacobi %d its %.3fs, "
        "multigrid %d its %.3fs" % (st["iterations"], t_l, st["fallback"],
                                    out["jacobi"]["iterations"], out["jacobi"]["time_s"],
                                    out["multigrid"]["iterations"], out["multigrid"]["time_s"]))
    tr.close()
    held.close()
    return out


def run_learned_warm(n_seeds=8, oc_train=40, steps_per_seed=500, eval_iters=12):
    """The learned warm start on the headline workload (120x60x60, multigrid flexible CG): the
    U-Net is trained on the warm-start residuals r = F - K u_prev of n_seeds other bracket-like
    problems (seeds 11..), one sample per OC iteration, steps_per_seed Adam steps after each
    problem; then the seed-2 problem runs eval_iters OC iterations with and without the learned
    step u += gamma M(r) before each warm multigrid solve. Reports wall time per OC iteration,
    FCG iterations, compliance agreement and the network's forward throughput on this grid."""
    import am_ml_suite_amd as A
    from am_ml_suite_amd.problems import make_config
    shape = (120, 60, 60)

    def handle(seed, iters):
        p = A.bracket_like_problem(*shape, seed=seed)
        o = A.TopologyOptimizer(p.nx, p.ny, p.nz, 1.0, A.MaterialProperties(1.0, 0.3),
                                make_config(p, preconditioner="multigrid", warm_start=True,
                                            max_iterations=iters))
        o.set_loads(p.F, p.fixed)
        o.set_density(p.rho0)
        o.begin_run()
        return o

    t0 = time.perf_counter()
    net, total = None, 0
    for s in range(n_seeds):
        o = handle(11 + s, oc_train)
        if net is None:
            net = A.UNet(optimizer=o, seed=3)
        o.collect_warm_samples(oc_train)
        for _ in range(oc_train):
            o.step()
        total += o.collect_warm_samples(oc_train)
        o.train_unet(net, steps_per_seed, lr=1e-3)
        o.close()
    t_train = time.perf_counter() - t0
    fwd_ms, fwd_tf = net.benchmark(10)
    runs = {}
    for key, use in (("multigrid", False), ("learned_multigrid", True)):
        o = handle(2, eval_iters)
        if use:
            o.attach_unet(net)
        o.sync()
        t0 = time.perf_counter()
        res = [o.step() for _ in range(eval_iters)]
        o.sync()
        dt = time.perf_counter() - t0
        runs[key] = dict(ms_per_oc=1e3 * dt / eval_iters,
                         fcg_iters=[r.solver_iterations for r in res],
                         max_true_rel_residual=max(r.true_rel_residual for r in res),
                         compliance=[r.compliance for r in res])
        o.close()
    cm = np.array(runs["multigrid"].pop("compliance"))
    cl = np.array(runs["learned_multigrid"].pop("compliance"))
    out = dict(workload="bracket_like_120x60x60_seed2", oc_iterations=eval_iters,
               train_problems=n_seeds, train_samples=total,
               train_steps=n_seeds * steps_per_seed, train_s=t_train,
               unet_forward_ms=fwd_ms, unet_conv_tflops=fwd_tf,
               unet_conv_frac_bf16_peak=fwd_tf / BF16_PEAK_TFLOPS,
               max_rel_compliance_diff=float(np.abs(cl / cm - 1).max()),
               speedup_learned_vs_multigrid=(runs["multigrid"]["ms_per_oc"]
                                             / runs["learned_multigrid"]["ms_per_oc"]),
               **runs)
    log("  [learned warm start] %d samples, %d steps, %.1fs; multigrid %.2f ms/OC (%.1f FCG its), "
        "learned+multigrid %.2f ms/OC (%.1f FCG its); U-Net forward %.3f ms, %.0f TFLOP/s"
        % (total, n_seeds * steps_per_seed, t_train, runs["multigrid"]["ms_per_oc"],
           np.mean(runs["multigrid"]["fcg_iters"]), runs["learned_multigrid"]["ms_per_oc"],
           np.mean(runs["learned_multigrid"]["fcg_iters"]), fwd_ms, fwd_tf))
    return out


def run_batch(warmup, steps, solver="jacobi"):
    import am_ml_suite_amd as A
    from am_ml_suite_amd.problems import make_config
    ps = A.batch_problems()
    kw = (dict(precision="mixed_fp32", warm_start=True) if solver == "jacobi"
          else dict(preconditioner="multigrid", warm_start=2))
    B = A.BatchOptimizer([(p.nx, p.ny, p.nz, 1.0, A.MaterialProperties(1.0, 0.3),
                           make_config(p, **kw)) for p in ps])
    for i, p in enumerate(ps):
        B.set_loads(i, p.F, p.fixed)
        B.set_density(i, p.rho0)
    if warmup:
        B.run(warmup)
    hist, el = B.run(steps)
    n_dof = sum(p.n_dofs for p in ps)
    flat = [r for row in hist for r in row]
    log("  [batch] %d OC iterations of %d problems in %.3fs" % (steps, len(ps), el))
    matvec = None
    if solver == "jacobi":  # the lockstep matvec the tuner picked, with every candidate's time
        cand, us = B.operator_info()
        tile, planes = A.BatchOptimizer.operator_shape(cand)
        matvec = dict(candidate=cand, tile=tile, planes=planes,
                      tuner_us={str(c): t for c, t in sorted(us.items())})
        log("  [batch] matvec: candidate %d (%s, %d planes), %.1f us; gather tile %.1f us"
            % (cand, tile, planes, us.get(cand, 0.0), us.get(0, 0.0)))
    return dict(workload="batch_problems_n64_seed3", n_problems=len(ps), n_dofs_total=n_dof,
                matvec=matvec,
                n_elements_total=sum(p.n_elements for p in ps),
                solver=("lockstep persistent launches, Jacobi-PCG, fp32 matvec + fp64"
                        if solver == "jacobi" else
                        "per-problem multigrid FCG (fp32 V-cycle, fp64 CG), 8 host threads"),
                steps=steps, warmup=warmup, time_s=el, oc_iter_per_s=steps / el,
                problem_oc_iter_per_s=steps * len(ps) / el, dof_oc_iter_per_s=n_dof * steps / el,
                max_pcg_iters=max(r.solver_iterations for r in flat),
                mean_pcg_iters=float(np.mean([r.solver_iterations for r in flat])),
                all_converged=all(r.solver_converged for r in flat),
                max_true_rel_residual=max(r.true_rel_residual for r in flat))


def cpu_baseline(problem):
    """Oracle (C++/OpenMP FP64 restatement of the reference path: assembled CSR per OC
    iteration, penalty BC, Jacobi-PCG, OpenMP sensitivities/filter/OC) on the host cores: one
    complete OC iteration of the 120x60x60 problem from its initial density, PCG uncapped to the
    1e-8 residual, timed per phase; and the 30x10x10 cantilever end to end, 50 OC iterations."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import am_ml_suite_amd as A
    import oracle_ffi as O
    p = problem
    cfg = O.make_config(p.nx, p.ny, p.nz, volume_fraction=p.params["volume_fraction"],
                        filter_radius=p.params["filter_radius"], max_iterations=1,
                        pcg_tol=1e-8, pcg_maxit=100000, ref_compat=0)
    t0 = time.perf_counter()
    _, hist, _, _ = O.run(cfg, p.rho0, p.F, p.fixed)
    wall = time.perf_counter() - t0
    h = hist[0]
    c1 = A.cantilever_problem()
    cfg1 = O.make_config(c1.nx, c1.ny, c1.nz, max_iterations=50, pcg_tol=1e-8, pcg_maxit=100000,
                         ref_compat=0)
    t0 = time.perf_counter()
    _, hist1, _, _ = O.run(cfg1, c1.rho0, c1.F, c1.fixed)
    wall1 = time.perf_counter() - t0
    return dict(value=1.0 / wall, unit="OC iter/s", cores=O.lib().orc_max_threads(),
                kind="port",
                sample=("oracle C++/OpenMP FP64 (assembled CSR per OC iteration, Jacobi-PCG to "
                        "1e-8, uncapped): the first OC iteration of the same 120x60x60 problem "
                        "from its initial density, timed whole (%d PCG passes, %.1f s); host %s"
                        % (h["solver_iterations"], wall,
                           platform.processor() or platform.machine())),
                pcg_iterations=h["solver_iterations"],
                phases_s=dict(assembly=h["assembly_time_s"], solve=h["solve_time_s"],
                              sensitivity=h["sensitivity_time_s"], update=h["update_time_s"]),
                cantilever_30x10x10=dict(oc_iterations=len(hist1), time_s=wall1,
                                         oc_iter_per_s=len(hist1) / wall1,
                                         pcg_iterations=int(sum(x["solver_iterations"]
                                                                for x in hist1))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--paths", default="plain,fast,learned,batch,simp_py,large,unet_conv")
    ap.add_argument("--large-steps", type=int, default=100)
    ap.add_argument("--fast-smoother", choices=("fp32", "mfma"), default="fp32",
                    help="fine-level smoothing of the fast path's V-cycle: the FP32 vector "
                         "operator (default) or the split-BF16 matrix-core operator")
    ap.add_argument("--operator-tile", type=int, default=-1,
                    help="level-0 operator launch of the fast path (amt_config.operator_tile): "
                         "-1 autotuned (default), or a fixed candidate (for counter runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (a plain copy kernel of "
                         "known size, to scale the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy << 20
        x = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "copies": 5}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import am_ml_suite_amd as A
    p = A.bracket_like_problem()
    log("problem %s: %d elements, %d DOFs, device %s"
        % (p.name, p.n_elements, p.n_dofs, A.device_name()))
    paths = args.paths.split(",")
    out = {}
    if "plain" in paths:
        log("plain FP64 Jacobi-PCG")
        out["plain"] = run_path(p, args.warmup, args.steps, "fp64", False, False)
    if "fast" in paths:
        log("fast: FP64 flexible CG + FP32 multigrid V-cycle + warm start")
        out["fast"] = run_path(p, args.warmup, args.steps,
                               "mixed_bf16" if args.fast_smoother == "mfma" else "fp64", 2,
                               True, "multigrid", args.operator_tile)

    large = None
    if "large" in paths:
        large = run_large(args.large_steps)
    learned = None
    if "learned" in paths:
        log("learned: U-Net warm start on the 120x60x60 workload")
        learned = dict(warm_start=run_learned_warm())
        log("learned: U-Net preconditioner on 96x48x48")
        learned["fcg_96x48x48"] = run_learned(train_steps=2000)
    unet_conv = None
    if "unet_conv" in paths:
        log("unet_conv: U-Net inference forward")
        unet_conv = run_unet_conv()
    ai_line = None
    if "ai" in paths:
        log("ai: density predictor")
        ai_line = run_ai()
    simp_py = None
    if "simp_py" in paths:
        log("simp_py: the Python SIMPOptimizer flavour")
        simp_py = run_simp_py(5, not args.no_cpu_baseline)
    batch = None
    if "batch" in paths or "batch_lockstep" in paths:
        log("batch: 64 generated problems in lockstep")
        batch = run_batch(1, 3)
        if "batch" in paths:  # batch_lockstep: the lockstep engine alone (one host thread)
            log("batch: the same 64 problems, multigrid per problem")
            batch["multigrid"] = run_batch(1, 3, "multigrid")

    if not out:  # only the auxiliary workloads were asked for
        line = {"metric": "auxiliary workloads", "batch": batch}
        if large is not None:
            line["large"] = large
        if simp_py is not None:
            line["simp_py"] = simp_py
        if learned is not None:
            line["learned"] = learned
        if ai_line is not None:
            line["ai"] = ai_line
        if unet_conv is not None:
            line["unet_conv"] = unet_conv
        print(json.dumps(line), flush=True)
        return

    main_key = "fast" if "fast" in out else "plain"
    m = out[main_key]
    ops = args.steps / m["time"]
    n_dof, n_e, n_n = p.n_dofs, p.n_elements, p.n_dofs // 3
    timed = m["results"][args.warmup:]
    pcg_per_oc = float(np.mean([r.solver_iterations for r in timed]))

    def op_name(base, op, tile):
        if op["candidate"] == 0:
            return "%s (%s)" % (base, tile)
        return ("%s_s<%d, %d> (element-column streaming tile, %d node planes)"
                % (base, op["tj"], op["tk"], op["planes"]))

    def roof(prof, bytes_per_launch, flops_per_launch, kernel, traffic):
        launches, ms = prof
        if not launches:
            return None
        avg_s = ms / launches / 1e3
        achieved = bytes_per_launch / avg_s / 1e9
        return dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS, traffic=traffic, kernel=kernel,
                    avg_launch_us=avg_s * 1e6, timed_launches=launches,
                    algorithmic_bytes_per_launch=bytes_per_launch,
                    flops_per_launch=flops_per_launch,
                    achieved_tflops=flops_per_launch / avg_s / 1e12)

    # The operator launch: p = z + beta p_old formed as the tile loads it, q = K p (FP64), p.q
    # partials; 33 B/DOF (read z, p_old; write p, q; the penalty count) + the 8-byte element
    # scale. flops: the dense 24x24 element product's 1152 per node.
    flops_op = 1152.0 * n_n
    roof_op = roof(m["prof"], n_dof * (4 * 8 + 1) + n_e * 8, flops_op,
                   "k_pcg_apply<double>" if main_key != "fast" else
                   op_name("k_fcg_apply", m["op"], "4x4x16 node gather tile"),
                   FCG_APPLY_TRAFFIC_BYTES.get(m["op"]["candidate"]) if main_key == "fast"
                   else None)
    roofline, kernels = roof_op, None
    if main_key == "fast":
        # The level-0 pre-smoothing with the flexible-CG update (one launch per FCG iteration,
        # like the operator): alpha from the operator's partials, r' = r - alpha q and
        # x += alpha p (FP64), x0 = wd r' and t = r' - K x0 (FP32 operator); 61 B/DOF (read r,
        # q, x, p, wd, the penalty count; write r', x, x0, t) + the 4-byte element scale.
        roof_upd = roof(m["prof_upd"], n_dof * 61 + n_e * 4, flops_op,
                        op_name("k_mg0_pre_upd", m["op_upd"], "4x4x16 node gather tile"),
                        PRE_UPD_TRAFFIC_BYTES.get(m["op_upd"]["candidate"]))
        kernels = dict(fcg_operator=roof_op, pre_smoothing_update=roof_upd)
        # the dominant kernel: the longer of the two launches of every FCG iteration
        if roof_upd and (roof_op is None or roof_upd["avg_launch_us"] > roof_op["avg_launch_us"]):
            roofline = roof_upd

    extra = {}
    if "plain" in out:
        pl = out["plain"]
        extra["plain_fp64_oc_iter_per_s"] = args.steps / pl["time"]
        extra["plain_pcg_iters_per_oc"] = float(np.mean(
            [r.solver_iterations for r in pl["results"][args.warmup:]]))
        if main_key == "fast":
            extra["speedup_vs_plain_fp64"] = ops / extra["plain_fp64_oc_iter_per_s"]
            cf = np.array([r.compliance for r in m["results"]])
            cp = np.array([r.compliance for r in pl["results"]])
            extra["max_rel_compliance_err_vs_fp64"] = float(np.abs(cf / cp - 1).max())
            extra["density_maxabs_vs_fp64"] = float(np.abs(m["density"] - pl["density"]).max())
    extra["fast_pcg_iters_per_oc"] = pcg_per_oc
    extra["dof_per_s"] = ops * n_dof
    extra["max_true_rel_residual"] = float(max(r.true_rel_residual for r in timed))

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle: one full OC iteration, and the cantilever)")
        cpu = cpu_baseline(p)

    line = {
        "metric": "OC iter/s at 1.3M DOF @1e-8 residual; learned+mixed vs plain FP64 PCG",
        "value": ops, "unit": "OC iter/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * m["time"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f64" if main_key != "fast" else
                  "bf16+f32+f64" if args.fast_smoother == "mfma" else "f32+f64"),
        "data": "synthetic",
        "config": {"workload": "bracket_like_120x60x60_seed2", "n_dofs": n_dof,
                   "n_elements": n_e, "pcg_tol": 1e-8, "path": main_key,
                   "solver": ("FP64 flexible CG + FP32 Galerkin multigrid V-cycle (Chebyshev "
                              "smoother on levels >= 2), warm start from the last two "
                              "displacements"
                              + (", fine-level smoothing on MFMA (split BF16)"
                                 if args.fast_smoother == "mfma" else "")
                              if main_key == "fast" else "FP64 Jacobi-PCG")},
        "roofline": roofline, "roofline_kernels": kernels, "cpu_baseline": cpu, "batch": batch,
    }
    if main_key == "fast":
        line["metric_note"] = (
            "value is the mixed-precision path without a network: FP64 flexible CG, FP32 "
            "Galerkin multigrid V-cycle, warm start from the energy-minimising combination of the "
            "previous two displacements. The learned "
            "part (U-Net warm start trained on the GPU) is timed on the same workload in "
            "learned.warm_start, against this path without it (speedup_learned_vs_multigrid), "
            "and is not in value. roofline is the longer of the two level-0 launches of an FCG "
            "iteration (roofline_kernels has both)")
    if large is not None:
        line["large"] = large
    if simp_py is not None:
        line["simp_py"] = simp_py
    if learned is not None:
        line["learned"] = learned
    if ai_line is not None:
        line["ai"] = ai_line
    if unet_conv is not None:
        line["unet_conv"] = unet_conv
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
