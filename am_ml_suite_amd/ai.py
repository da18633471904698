"""AI density predictor: the reference's AIOptimizer / TopOptUNet (src/am/ai/inference.py:19-120,
src/am/ai/model.py:18-67, placeholders that raise NotImplementedError) built on this package's
BF16 3D U-Net (hand-written HIP forward/backward, libamt.so) and trained on designs the GPU SIMP
engine produces.

Encoding (inference.py:72-90 names the channels: design mask, aggregated load directions, BC
positions), per element of the (nx, ny, nz) grid, zero-padded to the network's multiple-of-8 grid:
  0      design mask (grid == 0)
  1..3   load: the summed nodal forces of the element's 8 nodes over all load cases, / max |F|
  4..6   BC: fraction of the element's 8 nodes whose x / y / z DOF is constrained
  7      target volume fraction (constant)
  8, 9   fixed-solid and void masks
Output channel 0 is a density logit. Training minimises the binary cross-entropy between
sigmoid(logit) and the optimised density over the design elements (Adam on the host, gradients
from the device backward). Prediction shifts the logits so the design volume equals the target
(bisection on one offset); design elements holding a loaded node are set solid, fixed elements
are 1 and void ones 0.
"""
import numpy as np

from .optimizer import MaterialProperties, TopologyOptimizer, UNet
from .problems import DESIGN, FIXED, VOID, Problem, SplitMix64, _face_nodes, make_config

N_CH = 16


def _pad8(n):
    return -(-n // 8) * 8


def element_node_sum(nodal, nx, ny, nz):
    """(n_nodes, k) nodal values (node z-fastest) -> (nx, ny, nz, k) sums over each element's
    8 nodes."""
    a = nodal.reshape(nx + 1, ny + 1, nz + 1, -1)
    out = np.zeros((nx, ny, nz, a.shape[-1]))
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                out += a[di:di + nx, dj:dj + ny, dk:dk + nz]
    return out


def encode_problem(grid, load_cases, volume_fraction):
    """(16, nx, ny, nz) float32 input of the network (module docstring)."""
    grid = np.asarray(grid)
    nx, ny, nz = grid.shape
    n_dofs = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    F = np.zeros(n_dofs)
    cons = np.zeros(n_dofs)
    for lc in load_cases:
        F += lc.get_force_vector(n_dofs)
        cons[np.asarray(lc.get_constrained_dofs(), dtype=np.int64)] = 1.0
    x = np.zeros((N_CH, nx, ny, nz), dtype=np.float32)
    x[0] = grid == DESIGN
    Fe = element_node_sum(F.reshape(-1, 3), nx, ny, nz)
    fmax = np.abs(F).max()
    if fmax > 0:
        x[1:4] = np.moveaxis(Fe / fmax, -1, 0)
    x[4:7] = np.moveaxis(element_node_sum(cons.reshape(-1, 3), nx, ny, nz) / 8.0, -1, 0)
    x[7] = volume_fraction
    x[8] = grid == FIXED
    x[9] = grid == VOID
    return x


def bce_logits(z, t, mask):
    """Mean binary cross-entropy of sigmoid(z) against t over mask, and its gradient in z."""
    n = max(int(mask.sum()), 1)
    loss = (np.logaddexp(0.0, z) - t * z)[mask].sum() / n
    g = np.where(mask, (1.0 / (1.0 + np.exp(-z)) - t) / n, 0.0)
    return float(loss), g


class Adam:
    """Adam (Kingma & Ba) on a flat FP32 parameter vector."""

    def __init__(self, n, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n, np.float32)
        self.v = np.zeros(n, np.float32)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, p, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return (p - self.lr * mh / (np.sqrt(vh) + self.eps)).astype(np.float32)


def volume_matched(z, design, vf, iters=60):
    """sigmoid(z + b) with b chosen by bisection so the design elements' mean is vf."""
    lo, hi = -40.0, 40.0
    zd = z[design]
    for _ in range(iters):
        b = 0.5 * (lo + hi)
        if np.mean(1.0 / (1.0 + np.exp(-(zd + b)))) > vf:
            hi = b
        else:
            lo = b
    return 1.0 / (1.0 + np.exp(-(z + 0.5 * (lo + hi))))


class _LC:
    def __init__(self, F, fixed):
        self.F, self.fixed = F, fixed

    def get_force_vector(self, n_dofs):
        return self.F

    def get_constrained_dofs(self):
        return self.fixed


class TopOptUNet:
    """model.py:18-67: the density predictor's network on an (nx, ny, nz) element grid."""

    def __init__(self, shape, seed=1):
        self.shape = tuple(int(s) for s in shape)
        self.net = UNet(*(_pad8(s) for s in self.shape), seed=seed)

    def _pad(self, x):
        px, py, pz = self.net.shape
        nx, ny, nz = self.shape
        t = np.zeros((px, py, pz, N_CH), np.float32)
        t[:nx, :ny, :nz] = np.moveaxis(x, 0, -1)
        return t.reshape(-1, N_CH)

    def forward(self, x):
        """Density logits (nx, ny, nz) for an encoded input (16, nx, ny, nz)."""
        y = self.net.forward(self._pad(x)).reshape(*self.net.shape, N_CH)
        nx, ny, nz = self.shape
        return y[:nx, :ny, :nz, 0].astype(np.float64)

    def backward(self, dz):
        """Parameter gradient of sum(z * dz) for the last forward."""
        px, py, pz = self.net.shape
        nx, ny, nz = self.shape
        dy = np.zeros((px, py, pz, N_CH), np.float32)
        dy[:nx, :ny, :nz, 0] = dz
        return self.net.backward(dy.reshape(-1, N_CH))

    def predict(self, x):
        """model.py:55-67: sigmoid of the logits."""
        return 1.0 / (1.0 + np.exp(-self.forward(x)))


def ai_problems(n, shape=(24, 16, 16), seed=4):
    """Training/evaluation problems on one grid: all elements design, one random face clamped,
    1-3 point loads (unit direction, magnitude U(0.5, 1)) on random free nodes of the other
    faces, volume fraction U(0.2, 0.4), rmin 1.5; splitmix64 stream."""
    rng = SplitMix64(seed)
    nx, ny, nz = shape
    out = []
    for t in range(n):
        n_dofs = 3 * (nx + 1) * (ny + 1) * (nz + 1)
        clamp = rng.randint(0, 5)
        clamped = _face_nodes(nx, ny, nz, clamp)
        fixed = np.sort(np.concatenate([3 * clamped + c for c in range(3)])).astype(np.int64)
        F = np.zeros(n_dofs)
        others = [f for f in range(6) if f != clamp]
        for _ in range(rng.randint(1, 3)):
            nodes = np.setdiff1d(_face_nodes(nx, ny, nz, others[rng.randint(0, 4)]), clamped)
            node = int(nodes[rng.randint(0, len(nodes) - 1)])
            d = rng.unit_vector()
            mag = rng.uniform(0.5, 1.0)
            for c in range(3):
                F[3 * node + c] += mag * d[c]
        vf = rng.uniform(0.2, 0.4)
        params = dict(volume_fraction=vf, penalty=3.0, filter_radius=1.5, move_limit=0.2,
                      rho_min=0.01, E_min=1e-9)
        out.append(Problem("ai%d" % t, nx, ny, nz, np.full(nx * ny * nz, vf), F, fixed, params))
    return out


def to_grid(rho_slot, nx, ny, nz):
    """Slot order (x fastest) -> (nx, ny, nz)."""
    return np.asarray(rho_slot).reshape(nz, ny, nx).transpose(2, 1, 0)


def to_slot(rho_grid):
    return np.ascontiguousarray(np.asarray(rho_grid).transpose(2, 1, 0)).ravel()


def optimise(problems, iterations=40):
    """Ground-truth designs: the GPU SIMP engine (multigrid FCG, warm start) per problem.
    Returns the final densities as (nx, ny, nz) arrays."""
    out = []
    for p in problems:
        cfg = make_config(p, preconditioner="multigrid", warm_start=True,
                          max_iterations=iterations)
        opt = TopologyOptimizer(p.nx, p.ny, p.nz, 1.0, MaterialProperties(1.0, 0.3), cfg)
        opt.set_loads(p.F, p.fixed)
        opt.set_density(p.rho0)
        opt.begin_run()
        for _ in range(iterations):
            opt.step()
        out.append(to_grid(opt.get_density(), p.nx, p.ny, p.nz))
        opt.close()
    return out


def compliance_of(p, rho_grid):
    """c = F.u of the SIMP structure with density rho (GPU solve to 1e-8)."""
    cfg = make_config(p, preconditioner="multigrid")
    opt = TopologyOptimizer(p.nx, p.ny, p.nz, 1.0, MaterialProperties(1.0, 0.3), cfg)
    opt.set_loads(p.F, p.fixed)
    opt.set_density(to_slot(np.clip(rho_grid, p.params["rho_min"], 1.0)))
    u, _ = opt.solve()
    opt.close()
    return float(p.F @ u)


def refine(p, rho_grid, iterations):
    """Compliance after `iterations` OC iterations of the GPU SIMP engine started from
    rho_grid (clipped into the design range, so every element stays a design element:
    the engine's mask takes rho_min + 1e-6 < rho < 0.99, topology_optimizer.cpp:419-428)."""
    lo = p.params["rho_min"] + 1e-4
    cfg = make_config(p, preconditioner="multigrid", warm_start=True, max_iterations=iterations)
    opt = TopologyOptimizer(p.nx, p.ny, p.nz, 1.0, MaterialProperties(1.0, 0.3), cfg)
    opt.set_loads(p.F, p.fixed)
    opt.set_density(to_slot(np.clip(rho_grid, lo, 0.98)))
    opt.begin_run()
    for _ in range(iterations):
        opt.step()
    rho = to_grid(opt.get_density(), p.nx, p.ny, p.nz)
    opt.close()
    return compliance_of(p, rho)


class AIOptimizer:
    """inference.py:19-120 on the GPU: encode a problem, run the U-Net, post-process."""

    def __init__(self, model_path=None, device="cuda", shape=None, seed=1):
        self.device = device
        self.model = None
        if shape is not None:
            self.model = TopOptUNet(shape, seed=seed)
        if model_path is not None:
            self.load_model(model_path)

    def load_model(self, model_path):
        """Parameters from a .npz written by save_model (no pickles are loaded)."""
        with np.load(model_path, allow_pickle=False) as z:
            shape = tuple(int(v) for v in z["shape"])
            if self.model is None or self.model.shape != shape:
                self.model = TopOptUNet(shape)
            self.model.net.set_params(z["params"])

    def save_model(self, model_path):
        np.savez(model_path, shape=np.array(self.model.shape), params=self.model.net.get_params())

    def encode_problem(self, domain, load_cases, volume_fraction):
        return encode_problem(domain.grid, load_cases, volume_fraction)

    def predict_density(self, domain, load_cases, volume_fraction):
        """Continuous density with the target design volume. Design elements that carry a
        loaded node are made solid: a structure the network left short of a load point would
        otherwise hang that load on rho_min material."""
        grid = np.asarray(domain.grid)
        x = encode_problem(grid, load_cases, volume_fraction)
        z = self.model.forward(x)
        rho = volume_matched(z, grid == DESIGN, volume_fraction)
        loaded = np.abs(x[1:4]).sum(axis=0) > 0
        rho = np.where(loaded & (grid == DESIGN), 1.0, rho)
        return np.where(grid == FIXED, 1.0, np.where(grid == VOID, 0.0, rho))

    def predict(self, domain, load_cases, threshold=0.5, volume_fraction=0.3):
        """inference.py:92-109: the binarised predicted density (nx, ny, nz)."""
        rho = self.predict_density(domain, load_cases, volume_fraction)
        return (rho >= threshold).astype(np.float64)

    def train(self, samples, steps, lr=1e-3, batch=4, seed=0, log_every=0):
        """samples: list of (encoded input, target density, design mask). Returns the loss of
        every step (mean over its batch)."""
        p = self.model.net.get_params()
        opt = Adam(p.size, lr=lr)
        rng = np.random.default_rng(seed)
        losses = []
        for s in range(steps):
            g = np.zeros_like(p)
            tot = 0.0
            for i in rng.choice(len(samples), size=batch, replace=len(samples) < batch):
                x, t, m = samples[i]
                z = self.model.forward(x)
                loss, dz = bce_logits(z, t, m)
                g += self.model.backward(dz)
                tot += loss
            p = opt.step(p, g / batch)
            self.model.net.set_params(p)
            losses.append(tot / batch)
            if log_every and s % log_every == 0:
                print("  [ai] step %d loss %.4f" % (s, losses[-1]), flush=True)
        return losses

    def validate_against_numerical(self, problem, numerical_density, refine_iterations=0):
        """inference.py:111-120: IoU (threshold 0.5), volume error and compliance of the
        prediction against the optimised design; and the compliance after refine_iterations
        OC iterations started from the prediction, against the same number from the uniform
        density (the predictor as a warm start of the numerical optimiser)."""
        dom = _Domain(problem)
        lcs = [_LC(problem.F, problem.fixed)]
        vf = problem.params["volume_fraction"]
        rho = self.predict_density(dom, lcs, vf)
        a, b = rho >= 0.5, numerical_density >= 0.5
        iou = float((a & b).sum() / max((a | b).sum(), 1))
        c_pred = compliance_of(problem, rho)
        c_num = compliance_of(problem, numerical_density)
        out = dict(iou=iou, volume_error=float(abs((rho >= 0.5).mean() - vf)),
                   compliance=c_pred, compliance_numerical=c_num,
                   compliance_ratio=c_pred / c_num)
        if refine_iterations:
            out["refined_ratio"] = refine(problem, rho, refine_iterations) / c_num
            out["uniform_refined_ratio"] = refine(
                problem, np.full(rho.shape, vf), refine_iterations) / c_num
        return out


class _Domain:
    def __init__(self, p):
        self.grid = np.zeros((p.nx, p.ny, p.nz), dtype=np.int8)


def make_samples(problems, designs):
    out = []
    for p, d in zip(problems, designs):
        x = encode_problem(np.zeros((p.nx, p.ny, p.nz), np.int8), [_LC(p.F, p.fixed)],
                           p.params["volume_fraction"])
        out.append((x, d, np.ones(d.shape, dtype=bool)))
    return out
