"""Operator API: the drop-in for RadialBasisFunctions.jl's operator layer on this path.

Mirrors (names, argument meaning and error behaviour):
  AbstractOperator, RadialBasisOperator            src/operators/operators.jl:20-159
  apply op(x), op(y, x), mul!(y, op, x, a, b)      src/operators/operators.jl:173-186, 217-352, 529-542
  weights(op), update_weights!(op)                 src/operators/operators.jl:569-602
  laplacian / partial / mixed_partial              laplacian.jl:35-40, partial.jl:80-104, mixed_partial.jl:423-443
  jacobian / gradient / hessian                    jacobian.jl:218-257, gradient.jl:139-163, hessian.jl
  regrid / custom                                  regridding.jl:298-307, custom.jl:57-69
  Identity, ScaledOperator, SumOperator            operator_algebra.jl:83-280
  built-operator algebra (+, -, α*, /, unary -)    operator_algebra.jl:317-381
  find_neighbors, autoselect_k                     src/utils.jl:12-18, 39-43
An operator is lowered to a table of terms (kind, dims, coefficient, RHS column);
the weights are built on the GPU and kept there as CSR (one row per stencil).
"""
import ctypes as C

import numpy as np

from . import _abi as abi
from . import _lib
from . import functors
from .basis import AbstractRadialBasis, PHS, n_monomials, is_dirichlet
from .errors import ArgumentError, DimensionMismatch, SingularException

PRECISIONS = {"fp64": abi.PREC_FP64, "fp32_refined": abi.PREC_FP32_REFINED, "learned": abi.PREC_LEARNED}


# ============================================================== operator types
class AbstractOperator:
    """rank: tensor rank the operator adds to its output (operators.jl:20)."""

    rank = 0
    __array_ufunc__ = None  # numpy scalars defer to the operators below

    def lower(self, dim, basis=None):
        """Return (n_cols, [(col, kind, dim1, dim2, coeff), ...]) with 0-based dims; `basis`
        is the radial basis of the build (custom closures are lowered against it)."""
        raise NotImplementedError

    def __call__(self, data, **kw):
        return RadialBasisOperator(self, data, **kw)

    # symbolic algebra (operator_algebra.jl:180-232)
    def __add__(self, other):
        if isinstance(other, RadialBasisOperator):
            return NotImplemented
        if not isinstance(other, AbstractOperator):
            return NotImplemented
        if self.rank != other.rank:
            raise ArgumentError(
                f"Cannot combine operators with different output ranks: {print_op(self)} (rank {self.rank}) "
                f"and {print_op(other)} (rank {other.rank})")
        return SumOperator(_terms(self) + _terms(other))

    def __sub__(self, other):
        if not isinstance(other, AbstractOperator):
            return NotImplemented
        return self + ScaledOperator(-1, other)

    def __neg__(self):
        return ScaledOperator(-1, self)

    def __mul__(self, alpha):
        if isinstance(alpha, (int, float, np.floating, np.integer)):
            return ScaledOperator(alpha, self)
        return NotImplemented

    __rmul__ = __mul__

    def __truediv__(self, alpha):
        return ScaledOperator(1.0 / alpha, self)


def _terms(op):
    return list(op.ops) if isinstance(op, SumOperator) else [op]


class Laplacian(AbstractOperator):
    """∇² (laplacian.jl:6-7); for PHS the reference's 3-D constants are kept."""

    def lower(self, dim, basis=None):
        return 1, [(0, abi.OP_LAP, 0, 0, 1.0)]

    def __repr__(self):
        return "Laplacian (∇²f)"


class Partial(AbstractOperator):
    """∂ⁿ/∂x_dim, n ∈ {1, 2} (partial.jl:49-53, 92-104). dim is 1-based as in Julia."""

    def __init__(self, order, dim):
        self.order, self.dim = int(order), int(dim)

    def lower(self, dim, basis=None):
        if self.order not in (1, 2):
            raise ArgumentError("Only first and second order derivatives are supported. "
                                "Use the custom operator for higher orders.")
        if not 1 <= self.dim <= dim:
            raise ArgumentError(f"dimension {self.dim} out of range for {dim}-D data")
        kind = abi.OP_D1 if self.order == 1 else abi.OP_D2
        return 1, [(0, kind, self.dim - 1, 0, 1.0)]

    def __repr__(self):
        return f"∂ⁿf/∂xᵢ (n = {self.order}, i = {self.dim})"


class MixedPartial(AbstractOperator):
    """∂²/∂x_i∂x_j (mixed_partial.jl:388-400)."""

    def __init__(self, dim1, dim2):
        self.dim1, self.dim2 = int(dim1), int(dim2)

    def lower(self, dim, basis=None):
        for d in (self.dim1, self.dim2):
            if not 1 <= d <= dim:
                raise ArgumentError(f"dimension {d} out of range for {dim}-D data")
        if self.dim1 == self.dim2:
            return 1, [(0, abi.OP_D2, self.dim1 - 1, 0, 1.0)]
        return 1, [(0, abi.OP_MIXED, self.dim1 - 1, self.dim2 - 1, 1.0)]

    def __repr__(self):
        return f"∂²f/∂x{self.dim1}∂x{self.dim2}"


class Jacobian(AbstractOperator):
    """Tuple of first partials (jacobian.jl:184, operators.jl:33-35); rank 1."""

    rank = 1

    def __init__(self, dim):
        self.dim = int(dim)

    def lower(self, dim, basis=None):
        if dim != self.dim:
            raise ArgumentError(f"Jacobian{{{self.dim}}} applied to {dim}-D data")
        return dim, [(c, abi.OP_D1, c, 0, 1.0) for c in range(dim)]

    def __repr__(self):
        return "Jacobian (J)"


def _sym_pairs(D):
    """k-th (i, j) pair in upper-triangular row-major order (operators.jl:201-210)."""
    return [(i, j) for i in range(D) for j in range(i, D)]


class Hessian(AbstractOperator):
    """Upper-triangular second derivatives; rank 2."""

    rank = 2

    def __init__(self, dim):
        self.dim = int(dim)

    def lower(self, dim, basis=None):
        if dim != self.dim:
            raise ArgumentError(f"Hessian{{{self.dim}}} applied to {dim}-D data")
        return len(_sym_pairs(dim)), [
            (c, abi.OP_D2 if i == j else abi.OP_MIXED, i, j, 1.0) for c, (i, j) in enumerate(_sym_pairs(dim))]

    def __repr__(self):
        return "Hessian (H)"


class Regrid(AbstractOperator):
    """Interpolation to other points: Regrid()(b) = b (regridding.jl:266-267)."""

    def lower(self, dim, basis=None):
        return 1, [(0, abi.OP_VALUE, 0, 0, 1.0)]

    def __repr__(self):
        return "regrid"


class Identity(AbstractOperator):
    """Identity (operator_algebra.jl:83-85)."""

    def lower(self, dim, basis=None):
        return 1, [(0, abi.OP_VALUE, 0, 0, 1.0)]

    def __repr__(self):
        return "Identity (f)"


class ScaledOperator(AbstractOperator):
    """α · op (operator_algebra.jl:120-150)."""

    def __init__(self, alpha, op):
        self.alpha, self.op = alpha, op
        self.rank = op.rank

    @property
    def α(self):
        return self.alpha

    def lower(self, dim, basis=None):
        n, terms = self.op.lower(dim, basis)
        return n, [(c, k, d1, d2, self.alpha * w) for (c, k, d1, d2, w) in terms]

    def __repr__(self):
        return f"{self.alpha} × {print_op(self.op)}"


class SumOperator(AbstractOperator):
    """Sum of operators of one rank; nested sums flatten (operator_algebra.jl:195-208)."""

    def __init__(self, ops):
        self.ops = tuple(ops)
        self.rank = self.ops[0].rank

    def lower(self, dim, basis=None):
        lowered = [op.lower(dim, basis) for op in self.ops]
        ns = {n for n, _ in lowered}
        if len(ns) != 1:
            raise ArgumentError("Cannot sum gradient-family operators with different dimensions.")
        terms = [t for _, ts in lowered for t in ts]
        return ns.pop(), terms

    def __repr__(self):
        return " + ".join(print_op(o) for o in self.ops)


class Stacked(AbstractOperator):
    """Several operators sharing one stencil set, built as one system with one RHS
    column per output (no reference counterpart: the reference builds `gradient` and
    `laplacian` as two operators, i.e. two factorisations per stencil and two index
    arrays). Rank 1: applying it gives one output column per stacked column, in order
    (a Jacobian contributes `dim` columns)."""

    rank = 1

    def __init__(self, *ops):
        if not ops or not all(isinstance(o, AbstractOperator) for o in ops):
            raise ArgumentError("Stacked takes one or more operators")
        if any(o.rank > 1 for o in ops):
            raise ArgumentError("Stacked takes operators of rank 0 or 1")
        self.ops = ops

    def lower(self, dim, basis=None):
        n, terms = 0, []
        for o in self.ops:
            nc, t = o.lower(dim, basis)
            terms += [(c + n, kind, d1, d2, coeff) for c, kind, d1, d2, coeff in t]
            n += nc
        return n, terms

    def __repr__(self):
        return "Stacked(" + ", ".join(repr(o) for o in self.ops) + ")"


class Custom(AbstractOperator):
    """Custom{N}(ℒ) (custom.jl:10-14): ℒ maps a basis to a callable (x, xᵢ) -> value, or to
    a tuple of them for rank 1, written with basis(x, xᵢ) and the functors of
    `functors` (d, d2, dmixed, lap); an operator-algebra tree is accepted too.

    A closure is lowered against the build's basis to the device term table by tracing
    (functors.lower_closure); closures that are not constant-coefficient combinations
    of those terms raise ArgumentError. Calling a Custom on a basis returns ℒ(basis)."""

    def __init__(self, op, rank=None):
        if isinstance(op, AbstractOperator):
            self.op, self.fn = op, None
            self.rank = op.rank if rank is None else int(rank)
        elif callable(op):
            self.op, self.fn = None, op
            self.rank = functors.infer_rank(op) if rank is None else int(rank)
        else:
            raise ArgumentError(f"custom takes a closure basis -> (x, xi) -> value or an operator, "
                                f"got {type(op).__name__}")

    def __call__(self, data, **kw):
        if isinstance(data, AbstractRadialBasis):
            return self.fn(data) if self.fn is not None else self.op
        return RadialBasisOperator(self, data, **kw)

    def lower(self, dim, basis=None):
        if self.fn is None:
            return self.op.lower(dim, basis)
        return functors.lower_closure(self.fn, basis if basis is not None else PHS(3, poly_deg=2), dim, self.rank)

    def __repr__(self):
        return "Custom Operator"


def print_op(op):
    return repr(op)


# ============================================================== helpers
def autoselect_k(data, basis):
    """min(N, max(2·binomial(m+d, d), 2d+1)) (utils.jl:39-43)."""
    shape = tuple(data.shape) if hasattr(data, "shape") else np.asarray(data).shape
    N, d = shape[0], (shape[1] if len(shape) > 1 else 1)
    m = basis.poly_deg
    return min(N, max(2 * n_monomials(d, m) if m >= 0 else 0, 2 * d + 1))


def _float_type(x):
    """"float32" for single-precision points (a Float32 element type in the reference), else "float64"."""
    t = _lib.torch()
    if isinstance(x, t.Tensor):
        return "float32" if x.dtype == t.float32 else "float64"
    dt = getattr(x, "dtype", None)
    if dt is None and isinstance(x, (list, tuple)) and len(x):
        dt = getattr(x[0], "dtype", None)
    return "float32" if dt is not None and np.dtype(dt) == np.float32 else "float64"


def _as_points(x):
    t = _lib.torch()
    if isinstance(x, t.Tensor):
        x = x.detach()
        if x.ndim == 1:
            x = x[:, None]
        return x.to(dtype=t.float64).contiguous()
    a = np.asarray(x, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    return np.ascontiguousarray(a)


def _to_device(a, dev, dtype=None):
    t = _lib.torch()
    if isinstance(a, t.Tensor):
        return a.to(device=dev, dtype=dtype or a.dtype).contiguous()
    return t.as_tensor(np.ascontiguousarray(a), device=dev, dtype=dtype)


def find_neighbors(data, eval_points=None, k=None, device=None):
    """k nearest data points of each eval point, ascending, 0-based (utils.jl:12-18).

    Returns a numpy (N_eval, k) int32 array; the reference returns 1-based vectors."""
    if k is None:
        raise ArgumentError("k is required")
    dev = _lib.device_for(device)
    X = _to_device(_as_points(data), dev)
    Q = X if eval_points is None else _to_device(_as_points(eval_points), dev)
    return _lib.knn(X, Q, int(k)).cpu().numpy()


def reorder_points(x, *args, device=None):
    """reorder_points!(x, k) / reorder_points!(x, adjl, k) (utils.jl:45-55): permute the
    points of x in place into a locality-preserving order and return the permutation
    perm (0-based int64, new x = old x[perm]).

    The reference orders by reverse Cuthill-McKee on the k-NN graph; here the order is
    a Z-order (Morton) curve of the coordinates, computed on the GPU, which serves the
    same purpose: points that are close get nearby indices, so the X gathers of
    neighbouring stencils share cache lines. adjl and k keep the reference's call
    forms; the curve does not need them. x is a numpy array or torch tensor (N, d)."""
    if len(args) == 1:
        adjl, k = None, args[0]
    elif len(args) == 2:
        adjl, k = args
    else:
        raise TypeError("reorder_points(x, k) or reorder_points(x, adjl, k)")
    if isinstance(k, bool) or not isinstance(k, (int, np.integer)):
        raise TypeError(f"k must be an integer (got {type(k).__name__})")
    t = _lib.torch()
    is_torch = isinstance(x, t.Tensor)
    dev = x.device if (is_torch and x.is_cuda) else _lib.device_for(device)
    pts = _to_device(_as_points(x), dev)
    if adjl is not None and len(adjl) != pts.shape[0]:
        raise DimensionMismatch(f"adjl has {len(adjl)} entries, x has {pts.shape[0]} points")
    perm = _lib.spatial_order(pts).to(t.int64)
    if is_torch:
        x.copy_(x[perm.to(x.device)])
        return perm.cpu().numpy()
    perm = perm.cpu().numpy()
    x[...] = np.asarray(x)[perm]
    return perm


class Hermite:
    """Hermite boundary data: is_boundary (N,), bc (n_bnd BoundaryConditions),
    normals (n_bnd, d) -- the `hermite` NamedTuple of operators.jl:92-93."""

    def __init__(self, is_boundary, bc, normals):
        self.is_boundary, self.bc, self.normals = is_boundary, bc, normals


def _hermite_arrays(hermite, n_data, dim, dev):
    if isinstance(hermite, dict):
        isb, bcs, nrm = hermite["is_boundary"], hermite["bc"], hermite["normals"]
    elif isinstance(hermite, tuple):
        isb, bcs, nrm = hermite
    else:
        isb, bcs, nrm = hermite.is_boundary, hermite.bc, hermite.normals
    isb = np.asarray(isb, dtype=bool)
    nrm = np.asarray(nrm, dtype=np.float64).reshape(-1, dim) if len(nrm) else np.zeros((0, dim))
    # _validate_boundary_inputs (api.jl:399-419)
    if isb.shape[0] != n_data:
        raise DimensionMismatch(
            f"length(is_boundary) = {isb.shape[0]} must equal the number of data points ({n_data})")
    nb = int(isb.sum())
    if len(bcs) != nb or nrm.shape[0] != nb:
        raise DimensionMismatch(
            f"boundary_conditions ({len(bcs)}) and normals ({nrm.shape[0]}) must each have one entry per "
            f"boundary point (count(is_boundary) = {nb})")
    bidx = np.full(n_data, -1, dtype=np.int32)
    bidx[np.nonzero(isb)[0]] = np.arange(nb, dtype=np.int32)
    al = np.array([b.alpha for b in bcs], dtype=np.float64)
    be = np.array([b.beta for b in bcs], dtype=np.float64)
    t = _lib.torch()
    arrays = dict(
        is_boundary=t.as_tensor(isb.astype(np.uint8), device=dev),
        bnd_index=t.as_tensor(bidx, device=dev),
        bc_alpha=t.as_tensor(al if nb else np.zeros(1), device=dev),
        bc_beta=t.as_tensor(be if nb else np.zeros(1), device=dev),
        normals=t.as_tensor(nrm if nb else np.zeros((1, dim)), device=dev),
    )
    host = dict(is_boundary=isb, bnd_index=bidx, bc_alpha=al, bc_beta=be)
    return arrays, host, nb


def _adjacency(adjl, n_eval):
    """Accept an (N_eval, k) array, a list of index vectors (ragged -> CSR), or a CSR
    pair (ptr, idx) of numpy arrays or device tensors."""
    if isinstance(adjl, tuple) and len(adjl) == 2:
        t = _lib.torch()
        ptr, idx = adjl
        ptr_h = ptr.cpu().numpy() if isinstance(ptr, t.Tensor) else np.asarray(ptr)
        if ptr_h.shape[0] != n_eval + 1:
            raise DimensionMismatch(f"adjacency has {ptr_h.shape[0] - 1} stencils for {n_eval} eval points")
        lens = np.diff(ptr_h)
        idx = idx.to(t.int32) if isinstance(idx, t.Tensor) else np.ascontiguousarray(idx, dtype=np.int32)
        if lens.size and np.all(lens == lens[0]):
            return None, idx.reshape(n_eval, int(lens[0])), int(lens[0])
        return ptr_h.astype(np.int64), idx, int(lens.max()) if lens.size else 0
    if isinstance(adjl, np.ndarray) and adjl.ndim == 2:
        return None, np.ascontiguousarray(adjl, dtype=np.int32), adjl.shape[1]
    t = _lib.torch()
    if isinstance(adjl, t.Tensor) and adjl.ndim == 2:
        return None, adjl.to(t.int32).contiguous(), adjl.shape[1]
    lists = [np.asarray(a, dtype=np.int32).ravel() for a in adjl]
    if len(lists) != n_eval:
        raise DimensionMismatch(f"adjl has {len(lists)} stencils for {n_eval} eval points")
    lens = np.array([len(a) for a in lists], dtype=np.int64)
    if lens.size and np.all(lens == lens[0]):
        return None, np.stack(lists).astype(np.int32), int(lens[0])
    ptr = np.zeros(n_eval + 1, dtype=np.int64)
    np.cumsum(lens, out=ptr[1:])
    return ptr, np.concatenate(lists).astype(np.int32), int(lens.max())


# ============================================================== the operator
class RadialBasisOperator:
    """Operator of data using a radial basis with monomial augmentation
    (operators.jl:50-159). Weights live on the GPU as CSR; `weights()` returns
    scipy.sparse matrices (a tuple for rank >= 1)."""

    SPATIAL_ORDER_MIN = 65536  # below this the X vector sits in L2 anyway
    SPMV_MAPPING = 0  # rbfx_spmv_args.mapping: 0 automatic, 1 lanes per row, 2 staged through LDS, 3 vector loads,
    # 4 source tiles
    SOURCE_TILE_FIELDS = 8  # mapping 4 (source tiles) serves applies of this many fields or more

    def __init__(self, L, data, eval_points=None, basis=None, k=None, adjl=None, hermite=None, device=None,
                 precision="fp64", refine_steps=4, accept_tol=1e-12, model=None, poly_deg_map=None, _weights=None):
        """poly_deg_map (no reference counterpart): optional per-eval-point polynomial
        degree in -1..basis.poly_deg, e.g. a higher degree for stencils near a boundary;
        one build then mixes system sizes, which the library runs in size buckets."""
        if not isinstance(L, AbstractOperator):
            raise ArgumentError(f"{L!r} is not an operator")
        self.L = L
        self.basis = basis if basis is not None else PHS(3, poly_deg=2)
        if not isinstance(self.basis, AbstractRadialBasis):
            raise ArgumentError("basis must be a radial basis (PHS, IMQ or Gaussian)")
        self.dev = _lib.device_for(device)
        # Float32 points give Float32 weights and outputs (test/float32.jl:11-32): the system
        # is still assembled and solved in FP64, and the weights are rounded to FP32 once
        self.dtype = _float_type(data)
        self.data = _as_points(data)
        self.eval_points = self.data if eval_points is None else _as_points(eval_points)
        self.dim = self.data.shape[1]
        if self.eval_points.shape[1] != self.dim:
            raise DimensionMismatch("data and eval_points have different dimensions")
        self.precision = precision
        self.refine_steps, self.accept_tol = refine_steps, accept_tol
        self.model = model
        if precision == "learned" and model is None:
            raise ArgumentError('precision="learned" needs a trained model (learned.LearnedModel)')
        self.k = autoselect_k(self.data, self.basis) if k is None else int(k)
        self.hermite = hermite
        self._d_data = _to_device(self.data, self.dev)
        self._d_eval = self._d_data if eval_points is None else _to_device(self.eval_points, self.dev)
        t = _lib.torch()
        self._d_deg = None
        if poly_deg_map is not None:
            dm = np.asarray(poly_deg_map.cpu() if isinstance(poly_deg_map, t.Tensor) else poly_deg_map).ravel()
            if dm.shape[0] != self.eval_points.shape[0]:
                raise DimensionMismatch(f"poly_deg_map has {dm.shape[0]} entries for {self.eval_points.shape[0]} "
                                        "eval points")
            if dm.size and (dm.min() < -1 or dm.max() > self.basis.poly_deg):
                raise ArgumentError(f"poly_deg_map entries must lie in -1..{self.basis.poly_deg}")
            self._d_deg = t.as_tensor(dm.astype(np.int8), device=self.dev)
        if adjl is None:
            self.adj_ptr = None
            self.adj_idx = _lib.knn(self._d_data, self._d_eval, self.k)
        else:
            ptr, idx, kmax = _adjacency(adjl, self.n_eval)
            self.k = kmax
            self.adj_ptr = None if ptr is None else t.as_tensor(ptr, device=self.dev)
            self.adj_idx = _to_device(idx, self.dev, t.int32)
        self.status_counts = None
        self._host_weights = None
        self._row_order = None
        self._max_row = None
        self._plan = None
        self._tiles = None
        self._captured = False
        if _weights is None:
            self._build()
        else:
            self.row_ptr, self.col, self.vals, self.n_cols = _weights
        self.valid_cache = True

    # ---------------------------------------------------------------- sizes
    @property
    def n_eval(self):
        return self.eval_points.shape[0]

    @property
    def n_data(self):
        return self.data.shape[0]

    @property
    def rank(self):
        return self.L.rank

    # ---------------------------------------------------------------- build
    def _build(self):
        t = _lib.torch()
        n_cols, terms = self.L.lower(self.dim, self.basis)
        op = abi.make_operator(n_cols, terms)
        a = abi.BuildArgs()
        a.dim = self.dim
        a.k = self.k
        a.n_data = self.n_data
        a.n_eval = self.n_eval
        a.data = _lib.ptr(self._d_data)
        a.eval = _lib.ptr(self._d_eval)
        a.adj_ptr = _lib.ptr(self.adj_ptr)
        a.adj_idx = _lib.ptr(self.adj_idx)
        a.basis = self.basis.descriptor()
        a.op = op
        keep = []
        row_ptr = self.adj_ptr
        if self.hermite is not None:
            arrs, host, nb = _hermite_arrays(self.hermite, self.n_data, self.dim, self.dev)
            keep.append(arrs)
            a.is_boundary = _lib.ptr(arrs["is_boundary"])
            a.bnd_index = _lib.ptr(arrs["bnd_index"])
            a.bc_alpha = _lib.ptr(arrs["bc_alpha"])
            a.bc_beta = _lib.ptr(arrs["bc_beta"])
            a.normals = _lib.ptr(arrs["normals"])
            a.n_bnd = nb
            # exact row lengths (count_nonzeros, execution.jl:39-69): Dirichlet eval rows hold 1 entry
            lens = (np.diff(self.adj_ptr.cpu().numpy()) if self.adj_ptr is not None
                    else np.full(self.n_eval, self.k, dtype=np.int64))
            e_idx = np.arange(self.n_eval)
            isb = np.zeros(self.n_eval, dtype=bool)
            m = e_idx < self.n_data
            isb[m] = host["is_boundary"][e_idx[m]]
            dmask = np.zeros(self.n_eval, dtype=bool)
            if isb.any():
                bi = host["bnd_index"][e_idx[isb]]
                dmask[isb] = (host["bc_alpha"][bi] == 1.0) & (host["bc_beta"][bi] == 0.0)
            if dmask.any():
                lens = np.where(dmask, 1, lens)
                ptr = np.zeros(self.n_eval + 1, dtype=np.int64)
                np.cumsum(lens, out=ptr[1:])
                row_ptr = t.as_tensor(ptr, device=self.dev)
        nnz = int(row_ptr[-1].item()) if row_ptr is not None else self.n_eval * self.k
        # output buffers are reused across rebuilds (a captured graph keeps their addresses)
        old = getattr(self, "vals", None)
        if old is not None and tuple(old.shape) == (n_cols, nnz):
            vals = old
        else:
            vals = t.empty((n_cols, nnz), dtype=t.float64, device=self.dev)
        col = None
        if row_ptr is not None:
            oc = getattr(self, "col", None)
            col = oc if (oc is not None and oc.numel() == nnz and oc is not getattr(self, "_adj_flat", None)) \
                else t.empty(nnz, dtype=t.int32, device=self.dev)
        ost = getattr(self, "status", None)
        status = ost if (ost is not None and ost.numel() == self.n_eval) \
            else t.empty(self.n_eval, dtype=t.int32, device=self.dev)
        a.out_ptr = _lib.ptr(row_ptr)
        a.out_col = _lib.ptr(col)
        a.out_val = _lib.ptr(vals)
        a.out_nnz = nnz
        a.status = _lib.ptr(status)
        if self.precision not in PRECISIONS:
            raise ArgumentError(f"precision must be one of {sorted(PRECISIONS)}")
        a.precision = PRECISIONS[self.precision]
        a.refine_steps = int(self.refine_steps)
        a.accept_tol = float(self.accept_tol)
        a.stream = _lib.stream_ptr(self.dev)
        if self.precision == "learned":
            mstruct = self.model.struct(self.model._rho)
            keep.append(mstruct)
            a.model = C.addressof(mstruct)
        if self._row_order is None and self.n_eval >= self.SPATIAL_ORDER_MIN:
            # stencils solved (and rows applied) along a Z-curve of the eval points, so
            # neighbouring wavefronts gather neighbouring coordinates; results do not
            # depend on the order
            self._row_order = _lib.spatial_order(self._d_eval)
        a.order = _lib.ptr(self._row_order)
        a.poly_deg_map = _lib.ptr(self._d_deg)
        self._args, self._keep = a, keep
        rc, counts = _lib.build_weights(a, self.dev)
        self.status_counts = counts
        self.status = status
        if rc != 0:
            msg = _lib.last_error()
            if rc == abi.RBFX_ESINGULAR:
                raise SingularException(msg)
            if rc == abi.RBFX_EDIM:
                raise DimensionMismatch(msg)
            if rc == abi.RBFX_EUNSUPPORTED:
                raise NotImplementedError(msg)
            raise ArgumentError(msg)
        if self.dtype == "float32":
            vals.copy_(vals.to(t.float32))  # the applied weights are the FP32 weights
        self.row_ptr = row_ptr
        self._max_row = None
        self._adj_flat = self.adj_idx.reshape(-1)
        self.col = col if col is not None else self._adj_flat
        self.vals = vals
        self.n_cols = n_cols
        self._host_weights = None
        self._plan = None
        self._tiles = None

    def update_weights(self):
        """update_weights!(op) (operators.jl:589-602)."""
        self._build()
        self.valid_cache = True

    def _build_async(self):
        """Re-enqueue the last build on the current stream (no sync, no read-back)."""
        a = self._args
        a.stream = _lib.stream_ptr(self.dev)
        _lib.build_weights_async(a, self.dev)
        if self.dtype == "float32":
            self.vals.copy_(self.vals.to(_lib.torch().float32))

    def capture(self, x):
        """Record weight build + apply of the field x (n_data,) or (n_data, F) in a HIP
        graph. Returns a GraphedOperator whose replay() rebuilds the weights from the
        current coordinates and applies them to its `x` buffer into its `y` buffer
        (float64 buffers; a Float32 operator's replay applies its FP32-rounded weights)."""
        t = _lib.torch()
        xd = _to_device(x, self.dev, t.float64)
        X = (xd[:, None] if xd.ndim == 1 else xd.reshape(xd.shape[0], -1)).clone().contiguous()
        Y = t.empty((self.n_cols, self.n_eval, X.shape[1]), dtype=t.float64, device=self.dev)
        self._captured = True
        self._plan = None
        self._tiles = None
        self._build()               # blocking warm-up: sizes the library's lists
        self._spmv(X, out=Y)        # warm-up: computes the row order
        t.cuda.synchronize(self.dev)
        g = t.cuda.CUDAGraph()
        stream = t.cuda.Stream(self.dev)
        with t.cuda.graph(g, stream=stream):
            self._build_async()
            self._spmv(X, out=Y, asynchronous=True)
        return GraphedOperator(self, g, X, Y)

    def is_cache_valid(self):
        return self.valid_cache

    # ---------------------------------------------------------------- weights
    def weights(self):
        """Sparse weights: a scipy CSR matrix, or a tuple of them for rank >= 1."""
        import scipy.sparse as sp

        if not self.valid_cache:
            self.update_weights()
        if self._host_weights is None:
            ptr = (self.row_ptr.cpu().numpy() if self.row_ptr is not None
                   else np.arange(self.n_eval + 1, dtype=np.int64) * self.k)
            col = self.col.cpu().numpy()
            vals = self.vals.cpu().numpy().astype(self.dtype, copy=False)
            # each matrix owns its index arrays: scipy may canonicalise them in place
            mats = tuple(sp.csr_matrix((vals[c], col.copy(), ptr.copy()), shape=(self.n_eval, self.n_data))
                         for c in range(self.n_cols))
            self._host_weights = mats[0] if (self.rank == 0 and self.n_cols == 1) else mats
        return self._host_weights

    # ---------------------------------------------------------------- apply
    def _spmv(self, X, out=None, alpha=1.0, beta=0.0, asynchronous=False):
        """X: (n_data, F) float64 cuda, point-major. Returns (n_cols, n_eval, F)."""
        s, Y = self._spmv_args(X, out, alpha, beta, asynchronous)
        _lib.spmv(s, self.dev, asynchronous=asynchronous)
        return Y

    def _spmv_args(self, X, out=None, alpha=1.0, beta=0.0, asynchronous=False):
        """The rbfx_spmv arguments of Y = alpha * W X + beta * Y (and Y): a caller that
        applies the same operator to the same buffers many times can reuse them."""
        t = _lib.torch()
        F = X.shape[1]
        Y = out if out is not None else t.empty((self.n_cols, self.n_eval, F), dtype=t.float64, device=self.dev)
        s = abi.SpmvArgs()
        s.n_rows = self.n_eval
        s.n_cols = self.n_data
        s.row_ptr = _lib.ptr(self.row_ptr)
        s.k = self.k
        s.col_idx = _lib.ptr(self.col)
        s.vals = _lib.ptr(self.vals)
        s.nnz = self.vals.shape[1]
        s.n_valarrays = self.n_cols
        s.n_fields = F
        s.x = _lib.ptr(X)
        s.x_ld_point = X.stride(0)
        s.x_ld_field = X.stride(1)
        s.y = _lib.ptr(Y)
        s.y_ld_point = Y.stride(1)
        s.y_ld_field = Y.stride(2)
        s.y_ld_val = Y.stride(0)
        s.alpha = float(alpha)
        s.beta = float(beta)
        s.stream = _lib.stream_ptr(self.dev)
        if self._row_order is None and self.n_eval >= self.SPATIAL_ORDER_MIN:
            # rows walked along a Z-curve of the eval points: neighbouring rows gather
            # neighbouring X entries (the result does not depend on the order)
            self._row_order = _lib.spatial_order(self._d_eval)
        s.row_order = _lib.ptr(self._row_order)
        if self.row_ptr is not None and self._max_row is None:
            self._max_row = int((self.row_ptr[1:] - self.row_ptr[:-1]).max().item()) if self.n_eval > 0 else 0
        s.max_row_len = self._max_row or 0
        s.mapping = self.SPMV_MAPPING
        if self._row_order is not None and self.row_ptr is None and not asynchronous and not self._captured:
            # apply plan: the weights copied once into that order, so they stream
            # (not for captured operators: a graph replay rebuilds the weights in place)
            if self._plan is None:
                self._plan = _lib.permute_rows(self._row_order, self.k, self.col, self.vals)
            s.col_idx = _lib.ptr(self._plan[0])
            s.vals = _lib.ptr(self._plan[1])
            s.rows_permuted = 1
            if F >= self.SOURCE_TILE_FIELDS and self.n_cols == 1 and self.SPMV_MAPPING == 4:
                # source tiles: a tile's distinct X rows are staged in LDS once; 32 rows per
                # tile up to 16 fields (a Morton block of a 3-D regrid's 32 targets uses ~170
                # sources, 32 KB of 16 fields), fewer beyond, within the entries a tile
                # holds in LDS (1024 for one value array)
                rows = 32 if F <= 16 else (16 if F <= 32 else 8)
                while rows > 1 and rows * self.k > 1024:
                    rows //= 2
                if self._tiles is None or self._tiles[0] != rows:
                    self._tiles = (rows,) + _lib.source_tiles(self.n_eval, self.k, self._plan[0], rows)
                rows, tn, toff, tcol, tsrc = self._tiles
                s.tile_rows = rows
                s.tile_src = _lib.ptr(tsrc)
                s.tile_src_off = _lib.ptr(toff)
                s.tile_src_n = _lib.ptr(tn)
                s.tile_col = _lib.ptr(tcol)
        return s, Y

    def __call__(self, x, y=None):
        if not self.valid_cache:
            self.update_weights()
        t = _lib.torch()
        is_torch = isinstance(x, t.Tensor)
        xd = _to_device(x, self.dev, t.float64)
        if xd.shape[0] != self.n_data:
            raise DimensionMismatch(f"field has {xd.shape[0]} rows, operator has {self.n_data} data points")
        vec = xd.ndim == 1
        X = xd[:, None] if vec else xd.reshape(xd.shape[0], -1)
        Y = self._spmv(X)
        trailing = () if vec else tuple(xd.shape[1:])
        if self.rank == 0:
            res = Y[0].reshape((self.n_eval,) + trailing)
        elif self.rank == 1:
            res = Y.permute(1, 2, 0).reshape((self.n_eval,) + trailing + (self.n_cols,))
        else:
            D = self.dim
            full = t.empty((self.n_eval,) + trailing + (D, D), dtype=t.float64, device=self.dev)
            Yp = Y.reshape((self.n_cols, self.n_eval) + trailing)
            for c, (i, j) in enumerate(_sym_pairs(D)):
                full[..., i, j] = Yp[c]
                full[..., j, i] = Yp[c]
            res = full
        if self.dtype == "float32":
            res = res.to(t.float32)
        if y is not None:
            if isinstance(y, t.Tensor):
                y.copy_(res)
            else:
                y[...] = res.cpu().numpy()
            return y
        return res if is_torch else res.cpu().numpy()

    def mul(self, y, x, alpha=1.0, beta=0.0):
        """mul!(y, op, x, α, β) for rank-0 operators (operators.jl:529-542)."""
        if self.rank != 0:
            raise ArgumentError("mul! is defined for rank-0 operators with scalar weights")
        t = _lib.torch()
        xd = _to_device(x, self.dev, t.float64)
        X = xd[:, None] if xd.ndim == 1 else xd
        if isinstance(y, t.Tensor) and y.device == self.dev and y.dtype == t.float64 and y.is_contiguous():
            Y = y.view(1, self.n_eval, -1)
            self._spmv(X, out=Y, alpha=alpha, beta=beta)
            return y
        yd = _to_device(y, self.dev, t.float64)
        Y = yd.reshape(1, self.n_eval, -1).contiguous()
        self._spmv(X, out=Y, alpha=alpha, beta=beta)
        res = Y.reshape(yd.shape)
        if isinstance(y, t.Tensor):
            y.copy_(res)
        else:
            y[...] = res.cpu().numpy()
        return y

    @property
    def shape(self):
        return (self.n_eval, self.n_data)

    def __repr__(self):
        return (f"RadialBasisOperator\n├─Operator: {print_op(self.L)}\n├─Data type: {self.dtype}\n"
                f"├─Number of points: {self.n_data}\n├─Stencil size: {self.k}\n"
                f"└─Basis: {type(self.basis).__name__} with degree {self.basis.poly_deg} polynomial augmentation")

    # ---------------------------------------------------------------- algebra on built operators
    def _derive(self, L, vals):
        new = object.__new__(RadialBasisOperator)
        new.__dict__.update(self.__dict__)
        new.L = L
        if self.dtype == "float32":
            vals = vals.to(_lib.torch().float32).to(vals.dtype)
        new.vals = vals
        new._host_weights = None
        new._plan = None  # (the apply plan holds a copy of the values it was made from)
        new._tiles = None
        return new

    def _check_compatible(self, other):
        t = _lib.torch()
        if self.n_data != other.n_data or not np.allclose(self.data, other.data):
            raise ArgumentError("Can not add operators that were not built with the same data.")
        same = (self.col.shape == other.col.shape and t.equal(self.col, other.col)
                and ((self.row_ptr is None and other.row_ptr is None)
                     or (self.row_ptr is not None and other.row_ptr is not None
                         and t.equal(self.row_ptr, other.row_ptr))))
        if not same:
            raise ArgumentError("Can not add operators that do not have the same stencils.")

    def __add__(self, other):
        if not isinstance(other, RadialBasisOperator):
            return NotImplemented
        self._check_compatible(other)
        return self._derive(self.L + other.L, self.vals + other.vals)

    def __sub__(self, other):
        if not isinstance(other, RadialBasisOperator):
            return NotImplemented
        self._check_compatible(other)
        return self._derive(self.L - other.L, self.vals - other.vals)

    def __mul__(self, alpha):
        if isinstance(alpha, (int, float, np.floating, np.integer)):
            return self._derive(alpha * self.L, self.vals * alpha)
        return NotImplemented

    __rmul__ = __mul__

    def __truediv__(self, alpha):
        return self * (1.0 / alpha)

    def __neg__(self):
        return self * -1


def weights(op):
    return op.weights()


def update_weights(op):
    op.update_weights()


def is_cache_valid(op):
    return op.is_cache_valid()


# ============================================================== constructors
class GraphedOperator:
    """Build + apply of one operator recorded in a HIP graph (RadialBasisOperator.capture).
    `x` is the input field buffer (write new values into it), `y` the output
    (n_cols, n_eval, F); replay() re-runs the whole captured sequence."""

    def __init__(self, op, graph, x, y):
        self.op, self.graph, self.x, self.y = op, graph, x, y

    def replay(self):
        self.graph.replay()
        return self.y


def laplacian(data, basis=None, **kw):
    return RadialBasisOperator(Laplacian(), data, basis=basis, **kw)


def partial(data, order, dim, basis=None, **kw):
    return RadialBasisOperator(Partial(order, dim), data, basis=basis, **kw)


def mixed_partial(data, dim1, dim2, basis=None, **kw):
    return RadialBasisOperator(MixedPartial(dim1, dim2), data, basis=basis, **kw)


def jacobian(data, x=None, basis=None, **kw):
    pts = _as_points(data)
    op = RadialBasisOperator(Jacobian(pts.shape[1]), pts, basis=basis, **kw)
    return op if x is None else op(x)


def gradient(data, x=None, basis=None, **kw):
    return jacobian(data, x, basis=basis, **kw)


def hessian(data, basis=None, **kw):
    pts = _as_points(data)
    return RadialBasisOperator(Hessian(pts.shape[1]), pts, basis=basis, **kw)


def regrid(data, eval_points, basis=None, **kw):
    return RadialBasisOperator(Regrid(), data, eval_points=eval_points, basis=basis, **kw)


def custom(data, op, basis=None, rank=None, **kw):
    """custom(data, ℒ; rank, basis, ...) and custom(data, op::AbstractOperator) (custom.jl:57-69):
    ℒ a closure basis -> (x, xᵢ) -> value (rank inferred as in _infer_rank when omitted),
    or an operator-algebra tree, which is built as given (its rank is its own)."""
    if isinstance(op, AbstractOperator):
        return RadialBasisOperator(op, data, basis=basis, **kw)
    return RadialBasisOperator(Custom(op, rank), data, basis=basis, **kw)
