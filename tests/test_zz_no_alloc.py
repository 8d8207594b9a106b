"""The build that allocates nothing (rbfx_build_weights_noalloc), called through the C
ABI: two calls in a row, the second with a smaller input, each checked against
rbfx_build_weights on the same stencils with the same kernel (pinned, or the one the
blocking build tuned), equal to summation-order rounding, plus a ragged
adjacency list of one size class, and the limits the entry point refuses before doing
any work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
K_NSC = 5


def _args(RBF, gpu, pts, adj, basis, terms, n_cols=1, adj_ptr=None, precision=0):
    import torch

    A = RBF._abi
    keep = []
    d = torch.as_tensor(pts, device=gpu)
    if adj_ptr is None:
        n, k = adj.shape
        nnz = n * k
        aj = torch.as_tensor(adj.astype(np.int32).ravel(), device=gpu)
    else:
        n = len(adj_ptr) - 1
        k = int(np.diff(adj_ptr).max())
        nnz = int(adj_ptr[-1])
        aj = torch.as_tensor(adj.astype(np.int32), device=gpu)
        ap = torch.as_tensor(adj_ptr.astype(np.int64), device=gpu)
        keep.append(ap)
    vals = torch.full((n_cols * nnz,), np.nan, dtype=torch.float64, device=gpu)
    status = torch.full((n,), -1, dtype=torch.int32, device=gpu)
    keep += [d, aj, vals, status]
    a = A.BuildArgs()
    a.dim = pts.shape[1]
    a.k = k
    a.n_data = len(pts)
    a.n_eval = n
    a.data = d.data_ptr()
    a.eval = d.data_ptr()
    a.adj_idx = aj.data_ptr()
    if adj_ptr is not None:
        a.adj_ptr = keep[0].data_ptr()
        a.out_ptr = keep[0].data_ptr()
    a.basis = basis
    a.op = A.make_operator(n_cols, terms)
    a.out_val = vals.data_ptr()
    a.out_nnz = nnz
    a.status = status.data_ptr()
    a.precision = precision
    a.stream = RBF._lib.stream_ptr(gpu).value
    return a, keep, vals, status


def _noalloc(RBF, gpu, a):
    import torch

    rc = RBF._lib.lib().rbfx_build_weights_noalloc(C.byref(a))
    torch.cuda.synchronize()
    return rc


def _blocking(RBF, gpu, a):
    rc, _ = RBF._lib.build_weights(a, gpu)
    assert rc == 0, RBF._lib.last_error()


def _rows_close(w, ref, k, rtol):
    w, ref = w.reshape(-1, k), ref.reshape(-1, k)
    scale = np.abs(ref).max(1)
    return np.all(np.abs(w - ref).max(1) <= rtol * scale)


@pytest.mark.parametrize("pin", [True, False])
def test_noalloc_build_matches_blocking_build(RBF, gpu, pin):
    """300,000 then 40,000 stencils (2-D PHS3, degree 2, k = 20, Laplacian): the second,
    smaller call reuses the static area the first one filled."""
    A = RBF._abi
    basis = A.Basis(3, 2, 0.0)
    terms = [(0, A.OP_LAP, 0, 0, 1.0)]
    if pin:
        RBF._lib.tune_bucket(0, K_NSC)
    try:
        for n, seed in ((300_000, 7), (40_000, 8)):
            rng = np.random.default_rng(seed)
            pts = rng.random((n, 2))
            adj = RBF.find_neighbors(pts, k=20, device=gpu)
            # the blocking build first: unpinned, it tunes this configuration, and the
            # build without allocation then runs the kernel it picked
            b, keep2, vals2, _ = _args(RBF, gpu, pts, adj, basis, terms)
            _blocking(RBF, gpu, b)
            ref = vals2.cpu().numpy()
            a, keep, vals, status = _args(RBF, gpu, pts, adj, basis, terms)
            assert _noalloc(RBF, gpu, a) == 0, RBF._lib.last_error()
            w = vals.cpu().numpy()
            st = status.cpu().numpy()
            assert np.all(np.isfinite(w)) and np.all(st == A.ST_OK), np.unique(st)
            # one kernel on the same stencils; its offsets gathered in the kernel instead
            # of packed ahead (the same subtraction): equal to summation-order rounding
            assert _rows_close(w, ref, 20, 1e-13), n
    finally:
        RBF._lib.tune_bucket(0, -1)


def test_noalloc_ragged_one_size_class(RBF, gpu):
    """A ragged adjacency list with k + m <= 32 (k from 12 to 20, degree 2) in one launch."""
    from scipy.spatial import cKDTree

    A = RBF._abi
    rng = np.random.default_rng(9)
    n = 30_000
    pts = rng.random((n, 2))
    _, full = cKDTree(pts).query(pts, 20)
    ks = rng.integers(12, 21, size=n)
    ptr = np.concatenate([[0], np.cumsum(ks)])
    idx = np.concatenate([full[i, :ks[i]] for i in range(n)])
    basis = A.Basis(3, 2, 0.0)
    terms = [(0, A.OP_D1, 0, 0, 1.0)]
    b, keep2, vals2, _ = _args(RBF, gpu, pts, idx, basis, terms, adj_ptr=ptr)
    _blocking(RBF, gpu, b)
    ref = vals2.cpu().numpy()
    a, keep, vals, status = _args(RBF, gpu, pts, idx, basis, terms, adj_ptr=ptr)
    assert _noalloc(RBF, gpu, a) == 0, RBF._lib.last_error()
    w = vals.cpu().numpy()
    assert np.all(np.isfinite(w))
    for i in rng.choice(n, 2000, replace=False):
        s = slice(ptr[i], ptr[i + 1])
        assert np.abs(w[s] - ref[s]).max() <= 1e-10 * np.abs(ref[s]).max(), i


def test_noalloc_refuses_what_it_does_not_serve(RBF, gpu):
    """Checked on the host before any launch: FP32 + refinement, more stencils than the
    static area holds, and stencils of varying size above 32 rows."""
    A = RBF._abi
    pts = np.random.default_rng(1).random((64, 2))
    adj = np.tile(np.arange(20, dtype=np.int32), (64, 1))
    basis = A.Basis(3, 2, 0.0)
    terms = [(0, A.OP_LAP, 0, 0, 1.0)]
    a, keep, _, _ = _args(RBF, gpu, pts, adj, basis, terms, precision=A.PREC_FP32_REFINED)
    assert _noalloc(RBF, gpu, a) == A.RBFX_EUNSUPPORTED
    a, keep, _, _ = _args(RBF, gpu, pts, adj, basis, terms)
    a.n_eval = (1 << 22) + 1
    assert _noalloc(RBF, gpu, a) == A.RBFX_EINVAL
    assert "maximum" in RBF._lib.last_error()
    ptr = np.arange(0, 64 * 30 + 1, 30)
    a, keep, _, _ = _args(RBF, gpu, pts, np.tile(np.arange(30), 64), RBF._abi.Basis(3, 2, 0.0), terms, adj_ptr=ptr)
    assert _noalloc(RBF, gpu, a) == A.RBFX_EUNSUPPORTED
