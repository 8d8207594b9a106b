"""CPU oracle of the stencil-weight surrogate network (test infrastructure only).

Restates, in numpy float64, the forward pass, mean-squared-error loss, backward pass
and Adam update that the gfx950 kernels run (radialbasisfunctions.jl_amd/csrc/mlp.hip),
including where those kernels round to BF16: the network inputs, the BF16 weight copy,
each hidden activation after GELU, the loss gradient dY and every back-propagated dZ.
With `bf16=False` every rounding is skipped, which gives the exact float64 network
that the finite-difference checks pin.

The network has no counterpart in the reference package; it is pinned by
  - finite differences of its own float64 loss (tests/test_mlp_oracle.py),
  - the closed form of the tanh-GELU derivative, checked against differences.

Parameter layout (flat float32 array): per layer l, Wt_l [N_l x K_l] then b_l [N_l],
with (K, N) = (128, 256), (256, 256), (256, 256), (256, 32).
"""
import numpy as np

LAYERS = 4
IN_PAD, HID, OUT_PAD = 128, 256, 32
C_GELU = np.sqrt(2.0 / np.pi)


def k_of(l):
    return IN_PAD if l == 0 else HID


def n_of(l):
    return OUT_PAD if l == LAYERS - 1 else HID


def w_off(l):
    return sum(n_of(i) * k_of(i) + n_of(i) for i in range(l))


def b_off(l):
    return w_off(l) + n_of(l) * k_of(l)


N_PARAMS = w_off(LAYERS)


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    u2 = ((u + r) & 0xFFFF0000).astype(np.uint32)
    return u2.view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    u = C_GELU * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * 0.044715 * z * z)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    Ws = [p[w_off(l):w_off(l) + n_of(l) * k_of(l)].reshape(n_of(l), k_of(l)) for l in range(LAYERS)]
    bs = [p[b_off(l):b_off(l) + n_of(l)] for l in range(LAYERS)]
    return Ws, bs


def forward(params, X, bf16_round=True):
    """X: (M, 128) inputs. Returns (Y (M, 32), cache)."""
    rnd = bf16 if bf16_round else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, bs = unpack(params)
    Wq = [rnd(W) for W in Ws]
    H = rnd(X)
    Hs, Zs = [H], []
    for l in range(LAYERS):
        Z = Hs[-1] @ Wq[l].T + bs[l]
        Zs.append(Z)
        if l < LAYERS - 1:
            Hs.append(rnd(gelu(Z)))
    return Zs[-1], dict(Hs=Hs, Zs=Zs, Wq=Wq)


def loss_and_grad(params, X, T, n_out, bf16_round=True, row_weight=None):
    """Row-weighted mean squared error over the first n_out outputs,
    mean_m w_m mean_n (Y - T)^2, and its gradient (flat)."""
    rnd = bf16 if bf16_round else (lambda a: np.asarray(a, dtype=np.float64))
    Y, c = forward(params, X, bf16_round)
    M = X.shape[0]
    w = np.ones(M) if row_weight is None else np.asarray(row_weight, dtype=np.float64)
    D = np.zeros_like(Y)
    D[:, :n_out] = Y[:, :n_out] - T[:, :n_out]
    loss = float(np.mean(w[:, None] * D[:, :n_out] ** 2))
    dZ = rnd(2.0 * w[:, None] * D / (M * n_out))
    g = np.zeros(N_PARAMS)
    for l in range(LAYERS - 1, -1, -1):
        Hin = c["Hs"][l]
        g[w_off(l):w_off(l) + n_of(l) * k_of(l)] = (dZ.T @ Hin).ravel()
        g[b_off(l):b_off(l) + n_of(l)] = dZ.sum(0)
        if l > 0:
            dZ = rnd((dZ @ c["Wq"][l]) * gelu_grad(c["Zs"][l - 1]))
    return loss, g


def adam(p, g, m, v, lr, b1, b2, eps, t):
    """One Adam step (Kingma & Ba 2015, with bias correction); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def network_inputs(xi):
    """(n, k, 2) canonical offsets -> (n, 128) inputs: BF16 high parts of the 2k
    offsets, then the BF16 roundings of their remainders, then zeros."""
    n, k, _ = xi.shape
    flat = xi.reshape(n, 2 * k)
    hi = bf16(flat)
    lo = bf16(flat - hi)
    X = np.zeros((n, IN_PAD))
    X[:, :2 * k] = hi
    X[:, 2 * k:4 * k] = lo
    return X


def canonical_offsets(data, eval_points, adj):
    """Canonical radius-scaled offsets of 2-D stencils (the frame the network sees):
    xi = (x_i - x_e)/h, rotated so the first nonzero offset lies on +x, reflected
    so the rotated y offsets sum to >= 0 (quad-partial summation order as the kernel).
    Returns (xi, h, r1) with r1 the scaled distance of that first nonzero offset."""
    D = data[adj] - eval_points[:, None, :]
    n, k, _ = D.shape
    h = np.sqrt((D ** 2).sum(2).max(1))
    h[~(h > 0)] = 1.0
    X = D / h[:, None, None]
    nz = (X[..., 0] != 0) | (X[..., 1] != 0)
    first = np.argmax(nz, axis=1)
    ref = X[np.arange(n), first]
    r = np.sqrt(ref[:, 0] ** 2 + ref[:, 1] ** 2)
    c = np.where(r > 0, ref[:, 0] / np.where(r > 0, r, 1), 1.0)
    s = np.where(r > 0, ref[:, 1] / np.where(r > 0, r, 1), 0.0)
    xr = c[:, None] * X[..., 0] + s[:, None] * X[..., 1]
    yr = -s[:, None] * X[..., 0] + c[:, None] * X[..., 1]
    part = [yr[:, q::4].sum(1) for q in range(4)]
    tot = (part[0] + part[1]) + (part[2] + part[3])
    yr = np.where((tot < 0)[:, None], -yr, yr)
    return np.stack([xr, yr], 2), h, r
