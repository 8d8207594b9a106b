"""GPU tests of the learned weight path: the MFMA network kernels against the CPU
oracle (forward, gradients, Adam), the canonical-feature kernel, training targets
from the plain solver, and the learned build (acceptance, polynomial reproduction,
fallback parity, and the gain of the trained network over an untrained one).

Tolerances: BF16 operands with FP32 accumulation; the oracle rounds to BF16 at the
same points, so the remaining differences are FP32 summation order and an
occasional one-ulp BF16 rounding flip (relative Frobenius errors <= 2e-2)."""
import numpy as np
import pytest

import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _params(seed=0, scale=0.15):
    rng = np.random.default_rng(seed)
    p = np.zeros(MO.N_PARAMS, dtype=np.float32)
    for l in range(MO.LAYERS):
        n, k = MO.n_of(l), MO.k_of(l)
        fan_in = 80 if l == 0 else k
        W = np.zeros((n, k), dtype=np.float32)
        rows = 26 if l == MO.LAYERS - 1 else n
        W[:rows, :fan_in] = rng.uniform(-scale, scale, (rows, fan_in))
        p[MO.w_off(l):MO.w_off(l) + n * k] = W.ravel()
        p[MO.b_off(l):MO.b_off(l) + n][:rows] = rng.uniform(-0.1, 0.1, rows)
    return p


def _inputs(M, seed=1):
    rng = np.random.default_rng(seed)
    xi = rng.uniform(-1, 1, (M, 20, 2))
    return MO.network_inputs(xi)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _dev_model(RBF, gpu, p):
    import torch

    params = torch.as_tensor(p, device=gpu)
    wb = torch.empty(RBF._abi.MLP_N_WBF16, dtype=torch.bfloat16, device=gpu)
    RBF._lib.mlp_sync_bf16(params, wb)
    return params, wb


def test_mlp_forward_matches_oracle(RBF, gpu):
    import torch

    p = _params(0)
    for M in (1, 77, 256):
        X = _inputs(M, M)
        params, wb = _dev_model(RBF, gpu, p)
        x = torch.as_tensor(X, dtype=torch.float32, device=gpu).to(torch.bfloat16)
        y = torch.empty((M, 32), dtype=torch.float32, device=gpu)
        ws = torch.empty(RBF._lib.mlp_workspace_bytes(M), dtype=torch.uint8, device=gpu)
        RBF._lib.mlp_run(x, params, wb, ws, M, y_out=y)
        Y_ref, _ = MO.forward(p, X)
        Yg = y.cpu().numpy()
        assert _rel(Yg, Y_ref) < 2e-2, (M, _rel(Yg, Y_ref))
        assert np.max(np.abs(Yg - Y_ref)) < 5e-2 * np.max(np.abs(Y_ref))


def test_mlp_gradients_match_oracle(RBF, gpu):
    import torch

    M, n_out = 512, 26
    p = _params(2)
    X = _inputs(M, 5)
    rng = np.random.default_rng(7)
    T = np.zeros((M, 32))
    T[:, :n_out] = rng.standard_normal((M, n_out))
    params, wb = _dev_model(RBF, gpu, p)
    x = torch.as_tensor(X, dtype=torch.float32, device=gpu).to(torch.bfloat16)
    tgt = torch.as_tensor(T[:, :n_out].astype(np.float32), device=gpu).contiguous()
    grads = torch.empty(RBF._abi.MLP_N_PARAMS, dtype=torch.float32, device=gpu)
    loss = torch.zeros(1, dtype=torch.float32, device=gpu)
    ws = torch.empty(RBF._lib.mlp_workspace_bytes(M), dtype=torch.uint8, device=gpu)
    rw = rng.uniform(0.2, 3.0, M).astype(np.float32)
    RBF._lib.mlp_run(x, params, wb, ws, M, target=tgt, grads=grads, loss=loss, n_out=n_out,
                     row_weight=torch.as_tensor(rw, device=gpu))
    l_ref, g_ref = MO.loss_and_grad(p, X, T, n_out, row_weight=rw.astype(np.float64))
    g = grads.cpu().numpy().astype(np.float64)
    assert abs(loss.item() - l_ref) <= 1e-3 * l_ref
    for l in range(MO.LAYERS):
        n, k = MO.n_of(l), MO.k_of(l)
        gw, rw = g[MO.w_off(l):MO.w_off(l) + n * k], g_ref[MO.w_off(l):MO.w_off(l) + n * k]
        gb, rb = g[MO.b_off(l):MO.b_off(l) + n], g_ref[MO.b_off(l):MO.b_off(l) + n]
        assert _rel(gw, rw) < 2e-2, (l, "W", _rel(gw, rw))
        assert _rel(gb, rb) < 2e-2, (l, "b", _rel(gb, rb))


def test_adam_matches_oracle(RBF, gpu):
    import torch

    rng = np.random.default_rng(11)
    p = _params(4)
    g = rng.standard_normal(MO.N_PARAMS).astype(np.float32) * 1e-2
    params, wb = _dev_model(RBF, gpu, p)
    grads = torch.as_tensor(g, device=gpu)
    m = torch.zeros_like(grads)
    v = torch.zeros_like(grads)
    pr, mr, vr = p.astype(np.float64), np.zeros(MO.N_PARAMS), np.zeros(MO.N_PARAMS)
    for step in (1, 2, 3):
        RBF._lib.mlp_adam(params, grads, m, v, wb, 1e-3, 0.9, 0.999, 1e-8, step)
        pr, mr, vr = MO.adam(pr, g.astype(np.float64), mr, vr, 1e-3, 0.9, 0.999, 1e-8, step)
    assert np.max(np.abs(params.cpu().numpy() - pr)) < 1e-6
    # the BF16 copy follows the FP32 parameters
    wbn = wb.float().cpu().numpy()
    for l in range(MO.LAYERS):
        n, k = MO.n_of(l), MO.k_of(l)
        off = RBF._abi.mlp_wb_off(l)
        assert np.array_equal(wbn[off:off + n * k], MO.bf16(params.cpu().numpy()[MO.w_off(l):MO.w_off(l) + n * k]))


def test_features_match_oracle(RBF, O, gpu):
    import torch

    rng = np.random.default_rng(21)
    pts = rng.random((3000, 2))
    d = torch.as_tensor(pts, device=gpu)
    adj = RBF._lib.knn(d, d, 20)
    xi, x, h, r1 = RBF._lib.stencil_features(d, d, adj, 20)
    xi_ref, h_ref, r1_ref = MO.canonical_offsets(pts, pts, adj.cpu().numpy())
    assert np.max(np.abs(xi.cpu().numpy() - xi_ref)) < 1e-12
    assert np.max(np.abs(h.cpu().numpy() / h_ref - 1)) < 1e-14
    assert np.max(np.abs(r1.cpu().numpy() / r1_ref - 1)) < 1e-13
    xb = x.float().cpu().numpy()
    # inputs from the kernel's own offsets: BF16 high parts, remainders, zero padding
    assert np.array_equal(xb, MO.network_inputs(xi.cpu().numpy()))
    # the two parts carry the offsets to ~16 significant bits
    rec = xb[:, :40] + xb[:, 40:80]
    assert np.max(np.abs(rec - xi.cpu().numpy().reshape(-1, 40))) < 2e-5


def test_canonical_targets_match_oracle(RBF, O, gpu):
    """Training targets: the plain FP64 solve of canonical stencils, w~ and lambda."""
    import torch

    L = RBF.learned
    rng = np.random.default_rng(22)
    pts = rng.random((2000, 2))
    d = torch.as_tensor(pts, device=gpu)
    adj = RBF._lib.knn(d, d, 20)
    xi, _, h, _ = RBF._lib.stencil_features(d, d, adj, 20)
    w, lam = L.canonical_targets(xi, gpu)
    # the same stencils in the physical frame, oracle (reference arithmetic, long double)
    lap = O.abi.make_operator(1, [(0, O.abi.OP_LAP, 0, 0, 1.0)])
    _, _, wr, _, rc = O.build(pts, pts, adj.cpu().numpy(), O.basis_struct(3, 2), lap, scaled=True, long_double=True)
    assert rc == 0
    wt = w.cpu().numpy()
    wphys = wt / (h.cpu().numpy()[:, None] ** 2)
    wr = wr[0].reshape(-1, 20)
    err = np.abs(wphys - wr).max(1) / np.abs(wr).max(1)
    assert np.quantile(err, 0.99) <= 1e-10 and err.max() <= 1e-8
    # multipliers of the unit-radius system: the polynomial rows hold exactly
    assert np.all(np.isfinite(lam.cpu().numpy()))


def _ptw_residual(pts, adj, W):
    """max |P^T w - L P| / max|w| h^2 per stencil, Laplacian, degree 2 (physical frame)."""
    D = pts[adj] - pts[:, None, :]
    x, y = D[..., 0], D[..., 1]
    P = np.stack([np.ones_like(x), x, y, x * y, x * x, y * y], 2)
    LP = np.array([0, 0, 0, 0, 2, 2.0])
    h2 = (D ** 2).sum(2).max(1)
    r = np.einsum("nka,nk->na", P, W) - LP
    scale = np.abs(W).max(1) * h2
    return np.abs(r).max(1) / scale


@pytest.fixture(scope="module")
def trained(RBF, gpu):
    """A small model trained on the GPU (20k stencils, 1500 steps)."""
    L = RBF.learned
    X, Y, rw = L.training_set(20000, seed=99, device=gpu, clouds=2)
    model, hist = L.train(X, Y, steps=1500, batch=2048, lr=2e-3, seed=3, device=gpu, log=None, log_every=250,
                          row_weight=rw)
    return model, hist


def test_training_reduces_loss(trained):
    model, hist = trained
    assert hist[-1][1] < 0.5 * hist[0][1], hist


def test_learned_build_contract(RBF, O, gpu, trained):
    """Accepted stencils reproduce quadratics exactly (projection), have a collocation
    residual below the tolerance, and the rest are solver weights (parity with FP64)."""
    import torch

    model, _ = trained
    L = RBF.learned
    pts = L.density_cloud(6000, 4242)
    model.accept_tol = 0.05
    rho = torch.empty(len(pts), dtype=torch.float64, device=gpu)
    model._rho = rho
    op = RBF.laplacian(pts, basis=RBF.PHS(3, poly_deg=2), k=20, device=gpu, precision="learned", model=model)
    ref = RBF.laplacian(pts, basis=RBF.PHS(3, poly_deg=2), k=20, device=gpu, adjl=op.adj_idx, precision="fp64")
    st = op.status.cpu().numpy()
    acc = st == RBF._abi.ST_LEARNED
    assert acc.sum() + (st == RBF._abi.ST_REFINED).sum() + (st == RBF._abi.ST_FALLBACK64).sum() + \
        (st == RBF._abi.ST_FP64).sum() == len(pts)
    adj = op.adj_idx.cpu().numpy()
    W = op.vals.cpu().numpy()[0].reshape(-1, 20)
    Wr = ref.vals.cpu().numpy()[0].reshape(-1, 20)
    if acc.any():
        assert np.max(_ptw_residual(pts, adj, W)[acc]) < 1e-9
        assert np.all(rho.cpu().numpy()[acc] <= 0.05)
    rej = ~acc
    if rej.any():
        err = np.abs(W[rej] - Wr[rej]).max(1) / np.abs(Wr[rej]).max(1)
        assert err.max() < 1e-9
    model._rho = None


def test_trained_network_gains_over_untrained(RBF, gpu, trained):
    """What the trained network adds: at the same acceptance tolerance, the untrained
    network (same architecture, initial weights) gets essentially no stencil accepted,
    so the build falls back to the solver; the trained one accepts a share and its
    predicted weights are closer to the solver's."""
    import torch

    model, _ = trained
    L = RBF.learned
    Xv, Yv, _ = L.training_set(4000, seed=777, device=gpu, clouds=1)
    err_t = L.weight_errors(model, Xv, Yv)
    init = L.LearnedModel.initial(seed=3, device=gpu, out_mean=model.out_mean.cpu().numpy(),
                                  out_std=model.out_std.cpu().numpy())
    err_u = L.weight_errors(init, Xv, Yv)
    assert np.median(err_t) < 0.8 * np.median(err_u), (np.median(err_t), np.median(err_u))
    pts = L.density_cloud(4000, 31337)
    counts = {}
    for name, m in (("trained", model), ("untrained", init)):
        m.accept_tol = 0.2
        op = RBF.laplacian(pts, basis=RBF.PHS(3, poly_deg=2), k=20, device=gpu, precision="learned", model=m)
        counts[name] = op.status_counts[RBF._abi.ST_LEARNED]
    assert counts["trained"] > counts["untrained"], counts
