"""GPU test of the AI density predictor (am_ml_suite_amd.ai): the BF16 U-Net trained on designs
the GPU SIMP engine optimised, evaluated on held-out problems against the same network untrained
and against the uniform density of the same volume."""
import numpy as np
import pytest

from am_ml_suite_amd import ai

pytestmark = pytest.mark.gpu


def test_density_predictor_gain():
    """48 training + 8 held-out problems on a 16x8x8 grid (30 OC iterations each), 600 Adam
    steps: the training loss falls, and on the held-out problems the trained predictor's
    designs (same volume) are several times stiffer than the uniform density and than the
    untrained network's, with a higher IoU against the optimised designs."""
    shape = (16, 8, 8)
    probs = ai.ai_problems(56, shape=shape, seed=11)
    designs = ai.optimise(probs, 30)
    samples = ai.make_samples(probs, designs)
    A = ai.AIOptimizer(shape=shape, seed=1)
    U = ai.AIOptimizer(shape=shape, seed=1)
    losses = A.train(samples[:48], 600, lr=1e-3, batch=4, seed=0)
    first, last = np.mean(losses[:10]), np.mean(losses[-25:])
    out = {"trained": [], "untrained": [], "uniform": []}
    for p, d in zip(probs[48:], designs[48:]):
        t = A.validate_against_numerical(p, d)
        u = U.validate_against_numerical(p, d)
        out["trained"].append(t)
        out["untrained"].append(u)
        out["uniform"].append(ai.compliance_of(p, np.full(d.shape, p.params["volume_fraction"]))
                              / t["compliance_numerical"])
    r_t = np.median([r["compliance_ratio"] for r in out["trained"]])
    r_u = np.median([r["compliance_ratio"] for r in out["untrained"]])
    r_0 = np.median(out["uniform"])
    iou_t = np.mean([r["iou"] for r in out["trained"]])
    iou_u = np.mean([r["iou"] for r in out["untrained"]])
    print("loss %.3f -> %.3f; compliance / optimum (median): trained %.2f untrained %.2f "
          "uniform %.2f; IoU trained %.3f untrained %.3f" % (first, last, r_t, r_u, r_0, iou_t,
                                                            iou_u))
    assert last < 0.6 * first
    assert r_t < 0.3 * r_0 and r_t < 0.6 * r_u
    assert iou_t > iou_u + 0.15
    for p in probs[48:]:
        vf = p.params["volume_fraction"]
        rho = A.predict_density(ai._Domain(p), [ai._LC(p.F, p.fixed)], vf)
        assert abs(rho.mean() - vf) < 0.02  # volume matched; load-point elements add a little
