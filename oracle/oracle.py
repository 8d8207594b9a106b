"""ORACLE -- test infrastructure, not part of the product.

ctypes wrapper around oracle/liboracle.so, the CPU restatement of the
RadialBasisFunctions.jl weight build (see rbf_oracle.c for the file:line map).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.
"""
import ctypes as C
import importlib.util
import os
import pathlib

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent.parent
_LIB = None


def _abi():
    spec = importlib.util.spec_from_file_location(
        "_rbfx_abi_oracle", ROOT / "radialbasisfunctions.jl_amd" / "_abi.py"
    )
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


abi = _abi()


def lib():
    global _LIB
    if _LIB is None:
        path = ROOT / "oracle" / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
        L = C.CDLL(str(path))
        dp = C.POINTER(C.c_double)
        L.orc_phi.restype = C.c_double
        L.orc_phi.argtypes = [C.POINTER(abi.Basis), C.c_int, dp, dp]
        L.orc_lrbf.restype = C.c_double
        L.orc_lrbf.argtypes = [C.POINTER(abi.Basis), C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_lrbf_normal.restype = C.c_double
        L.orc_lrbf_normal.argtypes = [C.POINTER(abi.Basis), C.c_int, C.c_int, C.c_int, dp, dp, dp,
                                      C.POINTER(C.c_int)]
        L.orc_grad.argtypes = [C.POINTER(abi.Basis), C.c_int, dp, dp, dp]
        L.orc_hess.argtypes = [C.POINTER(abi.Basis), C.c_int, dp, dp, dp]
        L.orc_mon.argtypes = [C.c_int, C.c_int, dp, dp]
        L.orc_lmon.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_mon_normal.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.orc_n_monomials.restype = C.c_int
        L.orc_n_monomials.argtypes = [C.c_int, C.c_int]
        L.orc_sym_solve.restype = C.c_int
        L.orc_sym_solve.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.orc_sym_solve_ld.restype = C.c_int
        L.orc_sym_solve_ld.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.orc_build.restype = C.c_int
        L.orc_build.argtypes = [C.POINTER(abi.BuildArgs), C.c_int, C.c_int, C.c_int]
        L.orc_assemble.restype = C.c_int
        L.orc_assemble.argtypes = [C.POINTER(abi.BuildArgs), C.c_int64, C.c_int, dp]
        L.orc_knn.restype = C.c_int
        L.orc_knn.argtypes = [dp, C.c_int64, dp, C.c_int64, C.c_int, C.c_int,
                              C.POINTER(C.c_int32), C.c_int]
        L.orc_count_nonzeros.restype = C.c_int64
        L.orc_count_nonzeros.argtypes = [C.c_int64, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
        L.orc_spmv.argtypes = [C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32), dp, C.c_int,
                               dp, dp, C.c_int]
        for f in ("orc_sizeof_build_args", "orc_sizeof_spmv_args", "orc_sizeof_operator", "orc_sizeof_mlp_model",
                  "orc_sizeof_mlp_run_args"):
            getattr(L, f).restype = C.c_int64
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def basis_struct(kind, poly_deg=2, eps=1.0):
    return abi.Basis(int(kind), int(poly_deg), float(eps))


def phi(b, x, xi):
    x, xi = _f64(x), _f64(xi)
    return lib().orc_phi(C.byref(b), x.size, _dp(x), _dp(xi))


def lrbf(b, kind, x, xi, d1=0, d2=0):
    x, xi = _f64(x), _f64(xi)
    return lib().orc_lrbf(C.byref(b), x.size, kind, d1, d2, _dp(x), _dp(xi))


def lrbf_normal(b, kind, x, xi, n, d1=0):
    x, xi, n = _f64(x), _f64(xi), _f64(n)
    ok = C.c_int(0)
    v = lib().orc_lrbf_normal(C.byref(b), x.size, kind, d1, _dp(x), _dp(xi), _dp(n), C.byref(ok))
    return v if ok.value else None


def grad(b, x, xi):
    x, xi = _f64(x), _f64(xi)
    g = np.zeros(x.size)
    lib().orc_grad(C.byref(b), x.size, _dp(x), _dp(xi), _dp(g))
    return g


def hess(b, x, xi):
    x, xi = _f64(x), _f64(xi)
    h = np.zeros(x.size * x.size)
    lib().orc_hess(C.byref(b), x.size, _dp(x), _dp(xi), _dp(h))
    return h.reshape(x.size, x.size)


def n_monomials(dim, deg):
    return lib().orc_n_monomials(dim, deg)


def mon(dim, deg, x):
    x = _f64(x)
    out = np.zeros(max(n_monomials(dim, deg), 1))
    lib().orc_mon(dim, deg, _dp(x), _dp(out))
    return out


def lmon(dim, deg, kind, x, d1=0, d2=0):
    x = _f64(x)
    out = np.zeros(max(n_monomials(dim, deg), 1))
    lib().orc_lmon(dim, deg, kind, d1, d2, _dp(x), _dp(out))
    return out


def mon_normal(dim, deg, n, x):
    x, n = _f64(x), _f64(n)
    out = np.zeros(max(n_monomials(dim, deg), 1))
    lib().orc_mon_normal(dim, deg, _dp(n), _dp(x), _dp(out))
    return out


def sym_solve(A, B, long_double=False):
    A = np.array(A, dtype=np.float64, order="C")
    B = np.array(B, dtype=np.float64, order="C")
    vec = B.ndim == 1
    if vec:
        B = B[:, None].copy()
    n, r = B.shape
    X = np.zeros((n, r))
    if long_double:
        rc = lib().orc_sym_solve_ld(n, r, _dp(A), _dp(B), _dp(X))
    else:
        rc = lib().orc_sym_solve(n, r, _dp(A), _dp(B), _dp(X))
    if rc:
        raise np.linalg.LinAlgError("singular symmetric system")
    return X[:, 0] if vec else X


def knn(data, query, k, threads=0):
    data, query = _f64(data), _f64(query)
    dim = data.shape[1]
    out = np.zeros((query.shape[0], k), dtype=np.int32)
    rc = lib().orc_knn(_dp(data), data.shape[0], _dp(query), query.shape[0], dim, k,
                       out.ctypes.data_as(C.POINTER(C.c_int32)), threads)
    if rc:
        raise ValueError("orc_knn failed")
    return out


def _vp(a):
    return None if a is None else a.ctypes.data


def hermite_arrays(n_data, is_boundary, alpha, beta, normals):
    """Expand the reference's BoundaryData (types.jl:178-182) to the ABI arrays."""
    isb = np.ascontiguousarray(np.asarray(is_boundary, dtype=np.uint8))
    bidx = np.full(n_data, -1, dtype=np.int32)
    bidx[np.nonzero(isb)[0]] = np.arange(int(isb.sum()), dtype=np.int32)
    return (isb, bidx, _f64(alpha), _f64(beta), _f64(normals))


def build(data, eval_points, adj, basis, op, hermite=None, scaled=False, long_double=False,
          threads=0, _want_args=False):
    """Run the oracle build. adj: (n_eval, k) int array or (adj_ptr, adj_idx).
    op: abi.Operator. hermite: (is_boundary, alpha, beta, normals) per boundary point.
    Returns (out_ptr, out_col, out_val[n_cols, nnz], status, rc)."""
    data, ev = _f64(data), _f64(eval_points)
    n_data, dim = data.shape
    n_eval = ev.shape[0]
    a = abi.BuildArgs()
    a.dim = dim
    a.n_data = n_data
    a.n_eval = n_eval
    a.data = _vp(data)
    a.eval = _vp(ev)
    if isinstance(adj, tuple):
        adj_ptr = np.ascontiguousarray(adj[0], dtype=np.int64)
        adj_idx = np.ascontiguousarray(adj[1], dtype=np.int32)
        a.k = int(np.max(np.diff(adj_ptr))) if n_eval else 0
        a.adj_ptr = _vp(adj_ptr)
    else:
        adj_idx = np.ascontiguousarray(adj, dtype=np.int32)
        adj_ptr = None
        a.k = adj_idx.shape[1]
        a.adj_ptr = None
    a.adj_idx = _vp(adj_idx)
    a.basis = basis
    a.op = op
    keep = [data, ev, adj_idx, adj_ptr]
    isb = bidx = al = be = nr = None
    if hermite is not None:
        isb, bidx, al, be, nr = hermite_arrays(n_data, *hermite)
        keep += [isb, bidx, al, be, nr]
        a.is_boundary, a.bnd_index, a.bc_alpha, a.bc_beta, a.normals = map(_vp, (isb, bidx, al, be, nr))
        a.n_bnd = int(al.shape[0])
    out_ptr = np.zeros(n_eval + 1, dtype=np.int64)
    nnz = lib().orc_count_nonzeros(n_eval, _vp(adj_ptr), a.k, _vp(isb), _vp(bidx), _vp(al), _vp(be),
                                   out_ptr.ctypes.data_as(C.POINTER(C.c_int64)))
    out_col = np.zeros(nnz, dtype=np.int32)
    out_val = np.zeros((op.n_cols, nnz), dtype=np.float64)
    status = np.zeros(n_eval, dtype=np.int32)
    a.out_ptr = _vp(out_ptr)
    a.out_col = _vp(out_col)
    a.out_val = _vp(out_val)
    a.out_nnz = nnz
    a.status = _vp(status)
    if _want_args:
        return a, keep + [out_ptr, out_col, out_val, status]
    rc = lib().orc_build(C.byref(a), int(scaled), int(long_double), int(threads))
    return out_ptr, out_col, out_val, status, rc


def condition_numbers(data, eval_points, adj, basis, op, hermite=None, scaled=True):
    """2-norm condition number of each stencil's assembled saddle-point system (the
    matrix the solver factorises, in the radius-scaled frame by default); NaN for
    identity (Dirichlet) rows. Used to state per-stencil parity bounds."""
    a, keep = build(data, eval_points, adj, basis, op, hermite=hermite, scaled=scaled, _want_args=True)
    n_eval = a.n_eval
    m = lib().orc_n_monomials(a.dim, basis.poly_deg)
    nmax = (int(a.k) if not a.adj_ptr else int(np.max(np.diff(keep[3])))) + m
    buf = np.zeros(nmax * nmax)
    out = np.full(n_eval, np.nan)
    for e in range(n_eval):
        n = lib().orc_assemble(C.byref(a), e, int(scaled), _dp(buf))
        if n > 0:
            out[e] = np.linalg.cond(buf[:n * n].reshape(n, n))
    return out


def spmv(row_ptr, col, val, x, threads=0):
    x = _f64(x)
    vec = x.ndim == 1
    if vec:
        x = x[:, None].copy()
    F = x.shape[1]
    n = row_ptr.shape[0] - 1
    y = np.zeros((n, F))
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = _f64(val)
    lib().orc_spmv(n, row_ptr.ctypes.data_as(C.POINTER(C.c_int64)),
                   col.ctypes.data_as(C.POINTER(C.c_int32)), _dp(val), F, _dp(x), _dp(y), threads)
    return y[:, 0] if vec else y


def n_threads_default():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
