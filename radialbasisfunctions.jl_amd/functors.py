"""Basis-differentiation functors and the lowering of custom-operator closures.

The reference builds custom operators from a closure ℒ(basis) that returns a callable
`(x, xᵢ) -> value` written with the functors ∂, ∂², ∇, ∇², H and ∂mixed
(src/basis/basis.jl:21-92, src/operators/custom.jl:10-14, 57-69). Python identifiers
cannot be ∂ or ∇, so the functors here are

    d(basis, i)            ∂φ/∂xᵢ             (∂(basis, i))
    d2(basis, i)           ∂²φ/∂xᵢ²           (∂²(basis, i))
    dmixed(basis, i, j)    ∂²φ/∂xᵢ∂xⱼ         (∂mixed(basis, i, j))
    lap(basis)             ∇²φ                (∇²(basis))
    grad(basis)            ∇φ, a vector       (∇(basis))
    hess(basis)            the Hessian matrix (H(basis))

with 1-based dimensions as in Julia, and `basis(x, xᵢ)` for φ itself.

A closure is lowered by calling it once with two tracer points: `basis(x, xᵢ)` and the
functors return linear forms over the term kinds the device kernels assemble (value,
∂ᵢ, ∂ᵢ², ∂ᵢ∂ⱼ, ∇²), and sums, differences and scalings of them stay linear forms. Any
other use of the points (reading a coordinate, a product of two terms, a function of a
term) raises ArgumentError: the device term table holds constant coefficients. The
same term combination is applied to the monomial augmentation, as the reference's
functors do when one closure serves both bases (src/basis/basis.jl:13-19).

Called on ordinary coordinate arrays the functors evaluate the closed forms on the
host (src/basis/polyharmonic_spline.jl, inverse_multiquadric.jl, gaussian.jl); the
weights themselves are always built by the device kernels.
"""
import numbers

import numpy as np

from . import _abi as abi
from .basis import AbstractRadialBasis, AbstractPHS, IMQ, Gaussian, PHS
from .errors import ArgumentError

_NONLINEAR = ("custom closures are lowered to a constant-coefficient combination of basis "
              "derivatives (basis(x, xi), d, d2, dmixed, lap); {what} is not supported")


# ============================================================== tracing
class _TracerPoint:
    """Stands in for x or xᵢ while a closure is lowered; only functors may take it."""

    __array_ufunc__ = None  # numpy scalars defer to the reflected operators

    def __init__(self, name):
        self.name = name

    def _refuse(self, *a, **k):
        raise ArgumentError(_NONLINEAR.format(what=f"using the point {self.name} other than as a functor argument"))

    __getitem__ = __len__ = __iter__ = __array__ = __float__ = _refuse
    __add__ = __radd__ = __sub__ = __rsub__ = __mul__ = __rmul__ = __truediv__ = __rtruediv__ = _refuse
    __neg__ = __pow__ = __matmul__ = __rmatmul__ = _refuse


class LinearForm:
    """Σ coeff · term, term = (kind, dim1, dim2) with 0-based dims."""

    __array_ufunc__ = None  # numpy scalars defer to the reflected operators

    def __init__(self, coeffs):
        self.coeffs = dict(coeffs)

    @staticmethod
    def _scalar(a):
        return isinstance(a, numbers.Real) or (isinstance(a, np.ndarray) and a.ndim == 0)

    def __add__(self, other):
        if isinstance(other, LinearForm):
            c = dict(self.coeffs)
            for t, w in other.coeffs.items():
                c[t] = c.get(t, 0.0) + w
            return LinearForm(c)
        if self._scalar(other) and float(other) == 0.0:
            return self  # sum() starts from 0
        raise ArgumentError(_NONLINEAR.format(what="adding a constant to a basis term"))

    __radd__ = __add__

    def __sub__(self, other):
        return self + (-1.0) * other

    def __rsub__(self, other):
        return (-1.0) * self + other

    def __neg__(self):
        return (-1.0) * self

    def __mul__(self, a):
        if self._scalar(a):
            return LinearForm({t: float(a) * w for t, w in self.coeffs.items()})
        raise ArgumentError(_NONLINEAR.format(what="a product of two basis terms"))

    __rmul__ = __mul__

    def __truediv__(self, a):
        if self._scalar(a):
            return self * (1.0 / float(a))
        raise ArgumentError(_NONLINEAR.format(what="dividing by a basis term"))

    def _refuse(self, *a, **k):
        raise ArgumentError(_NONLINEAR.format(what="a nonlinear function of a basis term"))

    __pow__ = __rpow__ = __rtruediv__ = __float__ = __abs__ = __array__ = _refuse

    def terms(self):
        return [(k, d1, d2, w) for (k, d1, d2), w in self.coeffs.items() if w != 0.0]


# ============================================================== host closed forms
def _profile(basis, r2):
    """(φ, g1, g2): ∂ᵢφ = g1·Δᵢ and ∂ᵢ∂ⱼφ = g2·ΔᵢΔⱼ + δᵢⱼ·g1, Δ = x - xᵢ."""
    if isinstance(basis, AbstractPHS):
        n, r = basis.order, np.sqrt(r2)
        return r ** n, n * r ** (n - 2), n * (n - 2) * r ** (n - 4)
    e2 = float(basis.eps) ** 2
    if isinstance(basis, IMQ):
        s = e2 * r2 + 1.0
        return s ** -0.5, -e2 * s ** -1.5, 3.0 * e2 * e2 * s ** -2.5
    if isinstance(basis, Gaussian):
        phi = np.exp(-e2 * r2)
        return phi, -2.0 * e2 * phi, 4.0 * e2 * e2 * phi
    raise ArgumentError(f"no closed form for {basis!r}")


def evaluate_term(basis, kind, d1, d2, x, xi):
    """One term at the point pair (x, xᵢ) on the host; dims 0-based."""
    x, xi = np.asarray(x, dtype=np.float64).ravel(), np.asarray(xi, dtype=np.float64).ravel()
    D = x - xi
    r2 = float(D @ D)
    phi, g1, g2 = _profile(basis, r2)
    if kind == abi.OP_VALUE:
        return phi
    if kind == abi.OP_D1:
        return g1 * D[d1]
    if kind == abi.OP_D2:
        return g2 * D[d1] ** 2 + g1
    if kind == abi.OP_MIXED:
        return g2 * D[d1] * D[d2] + (g1 if d1 == d2 else 0.0)
    if kind == abi.OP_LAP:
        if isinstance(basis, AbstractPHS):
            # the reference's PHS Laplacians carry the 3-D constants n(n+1) r^(n-2) in every
            # dimension (polyharmonic_spline.jl:110-113, 196-199, 281-284, 367-370)
            n = basis.order
            return n * (n + 1) * np.sqrt(r2) ** (n - 2)
        return g2 * r2 + x.size * g1
    raise ArgumentError(f"unknown term kind {kind}")


# ============================================================== functors
class _Functor:
    def __init__(self, basis, kind, d1=0, d2=0):
        if not isinstance(basis, AbstractRadialBasis):
            raise ArgumentError("basis functors take a radial basis (PHS, IMQ or Gaussian)")
        self.basis, self.kind, self.d1, self.d2 = basis, kind, d1, d2

    def form(self):
        return LinearForm({(self.kind, self.d1, self.d2): 1.0})

    def __call__(self, x, xi):
        if isinstance(x, _TracerPoint) or isinstance(xi, _TracerPoint):
            _check_traced_basis(self.basis)
            return self.form()
        return evaluate_term(self.basis, self.kind, self.d1, self.d2, x, xi)


def _dim0(i):
    i = int(i)
    if i < 1:
        raise ArgumentError(f"dimensions are 1-based (got {i})")
    return i - 1


def d(basis, dim):
    """∂φ/∂x_dim (basis.jl:21-29)."""
    return _Functor(basis, abi.OP_D1, _dim0(dim))


def d2(basis, dim):
    """∂²φ/∂x_dim² (basis.jl:40-48)."""
    return _Functor(basis, abi.OP_D2, _dim0(dim))


def dmixed(basis, dim1, dim2):
    """∂²φ/∂x_dim1∂x_dim2 (basis.jl:69-92)."""
    i, j = _dim0(dim1), _dim0(dim2)
    if i == j:
        return _Functor(basis, abi.OP_D2, i)
    return _Functor(basis, abi.OP_MIXED, min(i, j), max(i, j))


def lap(basis):
    """∇²φ (basis.jl:50-56)."""
    return _Functor(basis, abi.OP_LAP)


class grad:
    """∇φ (basis.jl:31-38): a vector; in a closure, return a tuple of `d` functors instead."""

    def __init__(self, basis):
        self.basis = basis

    def __call__(self, x, xi):
        if isinstance(x, _TracerPoint) or isinstance(xi, _TracerPoint):
            raise ArgumentError("grad(basis) is vector-valued; a rank-1 custom closure returns "
                                "tuple(d(basis, i) for i in 1..dim)")
        n = np.asarray(x).size
        return np.array([evaluate_term(self.basis, abi.OP_D1, i, 0, x, xi) for i in range(n)])


class hess:
    """Hessian matrix of φ (basis.jl:58-67)."""

    def __init__(self, basis):
        self.basis = basis

    def __call__(self, x, xi):
        if isinstance(x, _TracerPoint) or isinstance(xi, _TracerPoint):
            raise ArgumentError("hess(basis) is matrix-valued; use dmixed(basis, i, j) terms in a closure")
        n = np.asarray(x).size
        return np.array([[evaluate_term(self.basis, abi.OP_MIXED, i, j, x, xi) for j in range(n)]
                         for i in range(n)])


# ============================================================== lowering
_TRACED = []


def _check_traced_basis(b):
    if _TRACED and b != _TRACED[-1]:
        def name(x):
            return f"{type(x).__name__}(poly_deg={x.poly_deg}, eps={x.eps})"

        raise ArgumentError(f"the closure uses {name(b)} but the operator's basis is {name(_TRACED[-1])}; "
                            "one build assembles one basis")


def basis_value(basis, x, xi):
    """basis(x, xᵢ): φ(‖x - xᵢ‖) (polyharmonic_spline.jl:50-51, inverse_multiquadric.jl:33-34,
    gaussian.jl:30-31), or the value term while a closure is traced."""
    if isinstance(x, _TracerPoint) or isinstance(xi, _TracerPoint):
        _check_traced_basis(basis)
        return LinearForm({(abi.OP_VALUE, 0, 0): 1.0})
    return evaluate_term(basis, abi.OP_VALUE, 0, 0, x, xi)


def _lower_item(item, basis):
    if isinstance(item, AbstractRadialBasis):   # the identity closure: basis -> basis
        _check_traced_basis(item)
        return [(abi.OP_VALUE, 0, 0, 1.0)]
    if isinstance(item, _Functor):
        _check_traced_basis(item.basis)
        return item.form().terms()
    if isinstance(item, LinearForm):
        return item.terms()
    if callable(item):
        try:
            out = item(_TracerPoint("x"), _TracerPoint("xi"))
        except TypeError as e:  # e.g. a numpy function of a term
            raise ArgumentError(_NONLINEAR.format(what=f"an operation that failed on a basis term ({e})")) from e
        if isinstance(out, LinearForm):
            return out.terms()
        if LinearForm._scalar(out) and float(out) == 0.0:
            return []
        raise ArgumentError(_NONLINEAR.format(what=f"a closure returning {type(out).__name__}"))
    raise ArgumentError(f"a custom closure must return a callable (x, xi) -> value or a tuple of them, "
                        f"got {type(item).__name__}")


def lower_closure(L, basis, dim, rank):
    """(n_cols, [(col, kind, d1, d2, coeff)]) of the closure ℒ applied to `basis`: one
    column for rank 0, one per tuple entry for rank 1."""
    _TRACED.append(basis)
    try:
        out = L(basis)
        is_tuple = isinstance(out, (tuple, list))
        if rank not in (0, 1) or is_tuple != (rank == 1):
            raise ArgumentError(f"a custom closure of rank {rank} must return "
                                + ("a tuple of callables, one per output column" if rank == 1
                                   else "one callable (x, xi) -> value"))
        items = list(out) if is_tuple else [out]
        terms = []
        for c, item in enumerate(items):
            for kind, d1, d2, w in _lower_item(item, basis):
                if max(d1, d2) >= dim:
                    raise ArgumentError(f"dimension {max(d1, d2) + 1} out of range for {dim}-D data")
                terms.append((c, kind, d1, d2, w))
    finally:
        _TRACED.pop()
    return len(items), terms


def infer_rank(L):
    """_infer_rank (custom.jl:16-25): 1 if ℒ(PHS(3; poly_deg=2)) is a tuple, else 0."""
    return 1 if isinstance(L(PHS(3, poly_deg=2)), (tuple, list)) else 0
