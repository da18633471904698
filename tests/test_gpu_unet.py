"""GPU tests of the BF16 MFMA U-Net against the FP64 numpy oracle, and of the learned
preconditioner pipeline (sample collection, on-GPU training, flexible CG with fallback)."""
import numpy as np
import pytest

import am_ml_suite_amd as A
from am_ml_suite_amd.problems import make_config
import unet_oracle as U

pytestmark = pytest.mark.gpu


def _net_and_oracle_inputs(seed=5, P=(16, 16, 16)):
    net = A.UNet(*P, seed=seed)
    params = net.get_params().astype(np.float64)
    # the device multiplies BF16 weights; biases stay FP32
    split = U.split(params)
    pb = np.concatenate([np.concatenate([U.bf16(W).ravel(), b]) for W, b in split])
    x = np.random.default_rng(seed).standard_normal(P + (16,))
    x[..., 4:] = 0.0
    return net, pb, x


# inputs of each conv layer, in the network's tensor names (concatenations for the skips)
_LAYER_IN = [("X0",), ("A0a",), ("P1",), ("A1a",), ("P2",), ("A2a",), ("P3",), ("B3a",),
             ("U2", "S2"), ("C2a",), ("U1", "S1"), ("C1a",), ("U0", "S0"), ("C0a",), ("C0",)]
_LAYER_OUT = ["A0a", "S0", "A1a", "S1", "A2a", "S2", "B3a", "B3", "C2a", "C2", "C1a", "C1",
              "C0a", "C0", "OUT"]


def _device_state(net):
    """The device's stored activations of the last forward, in the oracle's dict layout."""
    S = {t: net.activation(t).astype(np.float64) for t in A.UNet.TENSORS}
    for i, names in enumerate(_LAYER_IN):
        S["in%d" % i] = np.concatenate([S[t] for t in names], axis=-1)
    return S


# 16^3: one brick per block at every level, one input-channel group per split; 64 x 32 x 80
# (non-cubic): 640 level-0 bricks, so blocks walk several bricks, and the three input-channel
# groups of the first decoder layer run unsplit in one block (several groups per brick)
_SHAPES = [(16, 16, 16), (64, 32, 80)]


@pytest.mark.parametrize("P", _SHAPES)
def test_unet_forward_matches_oracle(P):
    """MFMA implicit-GEMM forward (BF16 storage, FP32 accumulation). Per layer, on the device's
    own input tensor: the stored output equals the FP64 conv (+ReLU) rounded to BF16 up to
    accumulation-order rounding flips (relative L2 <= 5e-3); pools and upsamplings likewise.
    End to end vs the oracle's BF16-rounded forward: relative L2 <= 3e-2."""
    net, pb, x = _net_and_oracle_inputs(P=P)
    y = net.forward(x.reshape(-1, 16)).reshape(x.shape)
    ref = U.forward(pb, x, rnd=U.bf16)['OUT']
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    S = _device_state(net)
    L = U.split(pb)
    errs = []
    for i, (_, _, _, _, relu) in enumerate(U.SPECS):
        z = U.conv(S["in%d" % i], *L[i])
        z = U.bf16(np.maximum(z, 0.0) if relu else z)
        out = S[_LAYER_OUT[i]]
        errs.append(np.linalg.norm(out - z) / np.linalg.norm(z))
    for a, b in (("P1", "S0"), ("P2", "S1"), ("P3", "S2")):
        # FP32 sum of eight BF16 values: exact unless their exponents differ widely
        np.testing.assert_allclose(S[a], U.bf16(U.pool(S[b])), rtol=2.0 ** -7, atol=0)
    for a, b in (("U2", "B3"), ("U1", "C2"), ("U0", "C1")):
        np.testing.assert_array_equal(S[a], U.up(S[b]))
    print("end-to-end rel L2 %.2e; per-layer:" % err, " ".join("%.1e" % e for e in errs))
    assert err <= 3e-2, err
    assert max(errs) <= 5e-3, errs


@pytest.mark.parametrize("P", _SHAPES)
def test_unet_parameter_gradients_match_oracle(P):
    """Backward (MFMA data gradient with flipped weights, FP32 weight gradient) vs FP64 numpy
    backprop through the device's stored forward state (same ReLU masks), with BF16 rounding
    at the device's storage points of activation gradients: per-layer relative L2 <= 2e-2.
    (Against the oracle's own forward the masks differ wherever a pre-activation lies within
    BF16 noise of zero; that alone moves these gradients by several percent.)"""
    net, pb, x = _net_and_oracle_inputs(seed=6, P=P)
    net.forward(x.reshape(-1, 16))
    dout = np.random.default_rng(7).standard_normal(x.shape)
    g = net.backward(dout.reshape(-1, 16)).astype(np.float64)
    gref = U.backward(pb, _device_state(net), U.bf16(dout), rnd=U.bf16)
    off, errs = 0, []
    for t, c1, c2, co, _ in U.SPECS:
        nw = t * co * (c1 + c2)
        sl = slice(off, off + nw + co)
        errs.append(np.linalg.norm(g[sl] - gref[sl]) / np.linalg.norm(gref[sl]))
        off += nw + co
    print("per-layer gradient rel L2:", " ".join("%.2e" % e for e in errs))
    assert max(errs) <= 2e-2, errs


def _bracket_opt(seed, prec, n=(40, 20, 20)):
    p = A.bracket_like_problem(*n, seed=seed)
    opt = A.TopologyOptimizer(p.nx, p.ny, p.nz, 1.0, A.MaterialProperties(1.0, 0.3),
                              make_config(p, preconditioner=prec))
    opt.set_loads(p.F, p.fixed)
    opt.set_density(p.rho0)
    return opt, p


def _eval_losses(opt, net, n):
    """Per-sample energies of M and of its Jacobi term on opt's samples (no parameter change)."""
    return opt.train_unet(net, n, lr=0.0, with_jacobi=True)


def test_learned_preconditioner_gain():
    """What the trained network gains over the method without it (Jacobi): collect Jacobi-PCG
    residuals on one bracket, train on the GPU, then on a held-out bracket (other loads and void)
    one application of M must remove more energy than the Jacobi term alone, on average, where
    the untrained network (gain 0) is exactly Jacobi. The learned flexible CG then solves the
    held-out problem to the same 1e-8 FP64 residual with compliance within 1e-6 of plain
    Jacobi-PCG (the fallback, reported, may take over)."""
    opt, p = _bracket_opt(4, "unet")
    held, ph = _bracket_opt(5, "unet")
    net = A.UNet(optimizer=opt, seed=3)
    assert net.gain == 0.0
    n = opt.collect_samples(64)
    nh = held.collect_samples(64)
    assert n >= 8 and nh >= 8
    l0, j0 = _eval_losses(held, net, nh)
    np.testing.assert_allclose(l0, j0, rtol=1e-12)  # gain 0: M is Jacobi
    losses, jl = opt.train_unet(net, 300, lr=1e-3, with_jacobi=True)
    assert np.all(np.isfinite(losses)) and np.all(np.isfinite(jl))
    adv = jl - losses  # energy the correction removes beyond the Jacobi term, per sample
    assert adv[-30:].mean() > max(adv[:30].mean(), 0.0)
    assert net.gain != 0.0
    lh, jh = _eval_losses(held, net, nh)
    print("held-out energy per application: learned %.3f vs Jacobi %.3f (gain %.3g)"
          % (lh.mean(), jh.mean(), net.gain))
    assert lh.mean() < jh.mean()
    held.attach_unet(net)
    held.set_density(ph.rho0)
    u, st = held.solve()
    assert st["converged"] and st["true_rel_residual"] <= 1.5e-8, st
    ref, _ = _bracket_opt(5, "jacobi")
    uref, sref = ref.solve()
    assert (u @ ph.F) == pytest.approx(uref @ ph.F, rel=1e-6)
    print("held-out solve: learned FCG %d iterations (fallback %s) vs Jacobi-PCG %d"
          % (st["iterations"], st["fallback"], sref["iterations"]))


@pytest.mark.parametrize("gain", [-1e6, 1e9])
def test_learned_preconditioner_forced_fallback(gain):
    """A corrupt correction weight makes M(r) = D^-1 r + gain * rms * net(.) indefinite (or
    dominated by an untrained network): the learned flexible CG stalls or loses positive
    curvature, and plain FP64 Jacobi-PCG takes over from the current iterate
    (linear_solver.cpp:148-152,175-180 report non-convergence and continue; here the host
    decides the hand-over). The fallback must be reported, reach the 1e-8 FP64 true residual and
    agree with plain Jacobi-PCG on the compliance within 1e-6; the reported iteration count
    includes the flexible-CG passes before the hand-over."""
    p = A.bracket_like_problem(40, 20, 20, seed=4)
    # a bounded iteration budget: a stall that the host failed to detect ends the test quickly
    opt = A.TopologyOptimizer(p.nx, p.ny, p.nz, 1.0, A.MaterialProperties(1.0, 0.3),
                              make_config(p, preconditioner="unet", solver_max_iterations=4000))
    opt.set_loads(p.F, p.fixed)
    opt.set_density(p.rho0)
    net = A.UNet(optimizer=opt, seed=7)
    net.gain = gain
    opt.attach_unet(net)
    u, st = opt.solve()
    ref, _ = _bracket_opt(4, "jacobi")
    uref, sref = ref.solve()
    print("gain %g: fallback %d, %d iterations (Jacobi-PCG alone %d), true residual %.2e"
          % (gain, st["fallback"], st["iterations"], sref["iterations"],
             st["true_rel_residual"]))
    assert st["fallback"] == 1
    assert st["converged"] and st["true_rel_residual"] <= 1.5e-8, st
    assert st["iterations"] > 0 and st["outer_iterations"] == 2
    assert (u @ p.F) == pytest.approx(uref @ p.F, rel=1e-6)
    opt.close()
    ref.close()
    net.close()


@pytest.mark.parametrize("shape", [(16, 16, 24), (24, 16, 40)])
def test_unet_inference_forward_matches_training_forward(shape):
    """The inference forward (decoder reading its upsampled inputs in place, split brick
    convolutions on the small levels) gives bitwise the training forward's output."""
    net = A.UNet(*shape, seed=5)
    x = np.random.default_rng(7).standard_normal((int(np.prod(shape)), 16)).astype(np.float32)
    y_train = net.forward(x)
    y_infer = net.infer(x)
    assert np.array_equal(y_train, y_infer)
    assert np.abs(y_train).max() > 0
    net.close()
