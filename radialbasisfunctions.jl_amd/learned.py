"""Learned stencil-weight surrogate: model container, training data and GPU training.

No counterpart in the reference package. A BF16 MLP maps the canonical, radius-scaled
stencil offsets (2k = 40 values, each given as a BF16 high part and a BF16 remainder)
through 256 -> 256 -> 256 (GELU) to the solution (w~, lambda) of the scaled
saddle-point system, multiplied by r1^2 (r1: the distance of the canonical reference
neighbour, which carries most of the weights' 1/r^2 scale). Training targets come from
the plain FP64 solver run on the canonical stencils themselves; the loss is the mean
squared error relative to each stencil's weight magnitude. Forward and backward passes
and the Adam update run in the library's own kernels (rbfx_mlp_run, rbfx_mlp_adam).
At build time (precision="learned") each prediction is projected onto exact
polynomial reproduction in FP64, checked by its collocation residual, and stencils
above `accept_tol` are re-solved by the FP32-refined solver.
"""
import ctypes as C
import math

import numpy as np

from . import _abi as abi
from . import _lib

DEFAULT_K = 20
N_MONO_DEG2_2D = 6


def density_cloud(n, seed, axis=0, flip=False):
    """n points in [0,1]^2 whose density grows x10 along `axis` (p(t) = (1+9t)/5.5)."""
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    t = (-1.0 + np.sqrt(1.0 + 99.0 * u)) / 9.0
    if flip:
        t = 1.0 - t
    o = rng.random(n)
    return np.stack([t, o], 1) if axis == 0 else np.stack([o, t], 1)


class LearnedModel:
    """Surrogate parameters resident on one device, plus the configuration it covers."""

    def __init__(self, params, out_mean, out_std, k=DEFAULT_K, accept_tol=1e-2, device=None):
        t = _lib.torch()
        dev = _lib.device_for(device)
        self.device = dev
        self.k = int(k)
        self.n_out = self.k + N_MONO_DEG2_2D
        self.params = t.as_tensor(np.asarray(params, dtype=np.float32), device=dev).contiguous()
        if self.params.numel() != abi.MLP_N_PARAMS:
            raise ValueError(f"expected {abi.MLP_N_PARAMS} parameters, got {self.params.numel()}")
        self.wbf16 = t.empty(abi.MLP_N_WBF16, dtype=t.bfloat16, device=dev)
        _lib.mlp_sync_bf16(self.params, self.wbf16)
        self.out_mean = t.as_tensor(np.asarray(out_mean, dtype=np.float64), device=dev).contiguous()
        self.out_std = t.as_tensor(np.asarray(out_std, dtype=np.float64), device=dev).contiguous()
        self.accept_tol = float(accept_tol)
        self.basis_kind = abi.RBFX_PHS3
        self.poly_deg = 2
        self.op_kind = abi.OP_LAP
        self._rho = None

    # -- construction -------------------------------------------------------------
    @classmethod
    def initial(cls, seed=0, k=DEFAULT_K, device=None, out_mean=None, out_std=None):
        """Glorot-uniform weights on the true fan-in (zero on the padded inputs/outputs)."""
        rng = np.random.default_rng(seed)
        p = np.zeros(abi.MLP_N_PARAMS, dtype=np.float32)
        n_out = k + N_MONO_DEG2_2D
        for l in range(abi.MLP_LAYERS):
            K, N = abi.mlp_k(l), abi.mlp_n(l)
            fan_in = 4 * k if l == 0 else K
            fan_out = n_out if l == abi.MLP_LAYERS - 1 else N
            a = math.sqrt(6.0 / (fan_in + fan_out))
            W = np.zeros((N, K), dtype=np.float32)
            W[:fan_out, :fan_in] = rng.uniform(-a, a, size=(fan_out, fan_in))
            p[abi.mlp_w_off(l):abi.mlp_w_off(l) + N * K] = W.ravel()
        om = np.zeros(n_out) if out_mean is None else out_mean
        osd = np.ones(n_out) if out_std is None else out_std
        return cls(p, om, osd, k=k, device=device)

    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy(), out_mean=self.out_mean.cpu().numpy(),
                 out_std=self.out_std.cpu().numpy(), k=self.k, accept_tol=self.accept_tol)

    @classmethod
    def load(cls, path, device=None):
        z = np.load(path, allow_pickle=False)
        return cls(z["params"], z["out_mean"], z["out_std"], k=int(z["k"]), accept_tol=float(z["accept_tol"]),
                   device=device)

    def struct(self, rho_out=None):
        m = abi.MlpModel()
        m.dim = 2
        m.k = self.k
        m.basis_kind = self.basis_kind
        m.poly_deg = self.poly_deg
        m.op_kind = self.op_kind
        m.n_out = self.n_out
        m.params = self.params.data_ptr()
        m.wbf16 = self.wbf16.data_ptr()
        m.out_mean = self.out_mean.data_ptr()
        m.out_std = self.out_std.data_ptr()
        m.accept_tol = self.accept_tol
        m.out_rho = None if rho_out is None else rho_out.data_ptr()
        return m

    # -- inference on prepared inputs (tests, calibration) ----------------------------
    def predict(self, X, batch=65536):
        """Network outputs (de-standardised, float64 (n, n_out)) for BF16 inputs X (n, 64)."""
        t = _lib.torch()
        n = X.shape[0]
        out = t.empty((n, abi.MLP_OUT_PAD), dtype=t.float32, device=self.device)
        ws = t.empty(_lib.mlp_workspace_bytes(min(batch, n)), dtype=t.uint8, device=self.device)
        for s in range(0, n, batch):
            m = min(batch, n - s)
            _lib.mlp_run(X[s:s + m], self.params, self.wbf16, ws, m, y_out=out[s:s + m])
        y = out[:, :self.n_out].double()
        return y * self.out_std + self.out_mean


# ---- training data ---------------------------------------------------------------
def canonical_targets(xi, device):
    """Plain FP64 solve of the canonical stencils: (w~ (n,k), lambda (n,6)) of the
    radius-scaled PHS3 + degree-2 Laplacian system."""
    t = _lib.torch()
    n, k, _ = xi.shape
    data = xi.reshape(n * k, 2).contiguous()
    ev = t.zeros((n, 2), dtype=t.float64, device=device)
    adj = t.arange(n * k, dtype=t.int32, device=device).reshape(n, k)
    out_val = t.empty(n * k, dtype=t.float64, device=device)
    out_lam = t.empty(n * N_MONO_DEG2_2D, dtype=t.float64, device=device)
    a = abi.BuildArgs()
    a.dim = 2
    a.k = k
    a.n_data = n * k
    a.n_eval = n
    a.data = data.data_ptr()
    a.eval = ev.data_ptr()
    a.adj_idx = adj.data_ptr()
    a.basis = abi.Basis(abi.RBFX_PHS3, 2, 0.0)
    a.op = abi.make_operator(1, [(0, abi.OP_LAP, 0, 0, 1.0)])
    a.out_val = out_val.data_ptr()
    a.out_nnz = n * k
    a.out_lambda = out_lam.data_ptr()
    a.precision = abi.PREC_FP64
    a.stream = _lib.stream_ptr(device).value
    rc, counts = _lib.build_weights(a, device)
    if rc != 0:
        raise _lib.RbfxError(rc, _lib.last_error())
    return out_val.reshape(n, k), out_lam.reshape(n, N_MONO_DEG2_2D)


def training_set(n_stencils, seed=1234, k=DEFAULT_K, device=None, clouds=4):
    """Inputs (BF16 (n,128)), targets (float64 (n, k+6): (w~, lambda) r1^2) and
    per-stencil loss weights (float32 (n,): 1 / mean(w~^2 r1^4), mean 1) from `clouds`
    seeded density clouds (x10 density variation along x, y, and reversed), each
    stencil in its canonical frame."""
    t = _lib.torch()
    dev = _lib.device_for(device)
    per = int(math.ceil(n_stencils / clouds))
    Xs, Ys = [], []
    for c in range(clouds):
        pts = density_cloud(per, seed + c, axis=c % 2, flip=(c // 2) % 2 == 1)
        d = t.as_tensor(pts, device=dev)
        adj = _lib.knn(d, d, k)
        xi, x, _, r1 = _lib.stencil_features(d, d, adj, k)
        w, lam = canonical_targets(xi, dev)
        Xs.append(x)
        Ys.append(t.cat([w, lam], 1) * (r1 * r1)[:, None])
        del d, adj, xi
    X = t.cat(Xs)[:n_stencils].contiguous()
    Y = t.cat(Ys)[:n_stencils].contiguous()
    rw = 1.0 / (Y[:, :k] ** 2).mean(1).clamp_min(1e-300)
    rw = (rw / rw.mean()).float().contiguous()
    return X, Y, rw


# ---- training ------------------------------------------------------------------
def train(X, Y, steps=12000, batch=4096, lr=1e-3, lr_min=2e-5, seed=0, k=DEFAULT_K, device=None, log_every=1000,
          model=None, log=print, row_weight=None):
    """Adam on the standardised targets (row-weighted squared error) with a cosine
    learning-rate schedule; every pass runs in the library's MFMA kernels.
    Returns (model, loss history)."""
    t = _lib.torch()
    dev = _lib.device_for(device)
    mean = Y.mean(0)
    std = Y.std(0).clamp_min(1e-12)
    if model is None:
        model = LearnedModel.initial(seed=seed, k=k, device=dev, out_mean=mean.cpu().numpy(),
                                     out_std=std.cpu().numpy())
    n_out = model.n_out
    Z = ((Y - mean) / std).float().contiguous()
    n = X.shape[0]
    ws = t.empty(_lib.mlp_workspace_bytes(batch), dtype=t.uint8, device=dev)
    grads = t.empty(abi.MLP_N_PARAMS, dtype=t.float32, device=dev)
    m = t.zeros_like(grads)
    v = t.zeros_like(grads)
    loss = t.zeros(1, dtype=t.float32, device=dev)
    g = t.Generator(device=dev)
    g.manual_seed(seed)
    perm = t.randperm(n, device=dev, generator=g)
    pos = 0
    hist = []
    for step in range(1, steps + 1):
        if pos + batch > n:
            perm = t.randperm(n, device=dev, generator=g)
            pos = 0
        idx = perm[pos:pos + batch]
        pos += batch
        xb = X[idx].contiguous()
        zb = Z[idx].contiguous()
        wb = None if row_weight is None else row_weight[idx].contiguous()
        _lib.mlp_run(xb, model.params, model.wbf16, ws, batch, target=zb, grads=grads, loss=loss, n_out=n_out,
                     row_weight=wb)
        cur = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * (step - 1) / steps))
        _lib.mlp_adam(model.params, grads, m, v, model.wbf16, cur, 0.9, 0.999, 1e-8, step)
        if step == 1 or step % log_every == 0 or step == steps:
            lv = float(loss.item())
            hist.append((step, lv))
            if log:
                log(f"  step {step:6d}  lr {cur:.2e}  loss {lv:.4e}")
    return model, hist


def weight_errors(model, X, Y):
    """Per-stencil relative max-norm error of the predicted weights against the
    solver's (both in the r1^2-scaled units of the targets; the ratio is scale-free)."""
    k = model.k
    P = model.predict(X)
    w_true = Y[:, :k]
    num = (P[:, :k] - w_true).abs().amax(1)
    den = w_true.abs().amax(1)
    return (num / den).cpu().numpy()


# ---- calibration of the acceptance tolerance -------------------------------------
def test_field(pts):
    """f = sin(4x) cos(3y) and its Laplacian -25 f."""
    f = np.sin(4 * pts[:, 0]) * np.cos(3 * pts[:, 1])
    return f, -25.0 * f


def calibrate(model, n=200_000, seed=97531, tols=(0.2, 0.1, 0.05, 0.03, 0.02, 0.01, 0.005, 0.002, 0.001),
              target_ratio=1.03, device=None, log=print):
    """Pick the largest acceptance tolerance whose applied Laplacian error on a held-out
    density cloud stays within `target_ratio` of the plain FP64 build's (f = sin4x cos3y).
    Sets model.accept_tol and returns the sweep [(tol, ratio, accepted fraction)]."""
    from . import operators as ops
    from .basis import PHS

    dev = _lib.device_for(device)
    pts = density_cloud(n, seed)
    f, ex = test_field(pts)
    plain = ops.laplacian(pts, basis=PHS(3, poly_deg=2), k=model.k, device=dev, precision="fp64")

    def nrmse(y):
        return float(np.sqrt(np.mean((y - ex) ** 2)) / np.sqrt(np.mean(ex ** 2)))

    e0 = nrmse(plain(f))
    sweep = []
    chosen = None
    for tol in sorted(tols, reverse=True):
        model.accept_tol = tol
        op = ops.laplacian(pts, basis=PHS(3, poly_deg=2), k=model.k, device=dev, adjl=plain.adj_idx,
                           precision="learned", model=model)
        ratio = nrmse(op(f)) / e0
        acc = op.status_counts[abi.ST_LEARNED] / n
        sweep.append((tol, ratio, acc))
        if log:
            log(f"  tol {tol:.3g}: error ratio {ratio:.4f}, accepted {acc:.3f}")
        if chosen is None and ratio <= target_ratio:
            chosen = tol
    model.accept_tol = chosen if chosen is not None else min(tols)
    return sweep


def train_default(device=None, n_train=2_000_000, steps=20000, batch=4096, seed=2024, log=print):
    """The 2-D Laplacian surrogate (PHS3, degree 2, k=20): 2M training stencils from
    seeded density clouds, Adam with a cosine schedule, then tolerance calibration."""
    dev = _lib.device_for(device)
    X, Y, rw = training_set(n_train, seed=seed, device=dev)
    model, hist = train(X, Y, steps=steps, batch=batch, seed=seed, device=dev, log=log, row_weight=rw)
    del X, Y, rw
    sweep = calibrate(model, device=dev, log=log)
    return model, hist, sweep
