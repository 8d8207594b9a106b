"""Radial bases, monomial bases and boundary conditions (host-side descriptions).

Mirrors the constructors and argument checks of the reference:
  PHS / PHS1 / PHS3 / PHS5 / PHS7   src/basis/polyharmonic_spline.jl:113-126, 135-141, 473-477
  IMQ                               src/basis/inverse_multiquadric.jl:15-31
  Gaussian                          src/basis/gaussian.jl:12-20
  MonomialBasis                     src/basis/monomial.jl:6-16
  check_poly_deg                    src/utils.jl:57-66
  BoundaryCondition and predicates  src/solve/types.jl:18-42
The basis arithmetic itself runs in the device kernels.
"""
from math import comb

from . import _abi as abi
from .errors import ArgumentError


def check_poly_deg(poly_deg):
    if poly_deg < -1:
        raise ArgumentError(
            f"poly_deg must be >= -1 (got {poly_deg}). Use poly_deg=2 (default) for quadratic, "
            "poly_deg=0 for constant, or poly_deg=-1 to disable polynomial augmentation."
        )


class AbstractRadialBasis:
    kind = None
    eps = 1.0

    def __init__(self, poly_deg=2):
        poly_deg = int(poly_deg)
        check_poly_deg(poly_deg)
        self.poly_deg = poly_deg

    def descriptor(self):
        return abi.Basis(self.kind, self.poly_deg, float(self.eps))

    def __call__(self, x, xi):
        """φ(‖x - xᵢ‖) on the host (polyharmonic_spline.jl:50-51, inverse_multiquadric.jl:33-34,
        gaussian.jl:30-31); inside a custom closure being lowered, the value term."""
        from .functors import basis_value

        return basis_value(self, x, xi)

    def __eq__(self, other):
        return type(self) is type(other) and self.poly_deg == other.poly_deg and self.eps == other.eps

    def __hash__(self):
        return hash((type(self).__name__, self.poly_deg, self.eps))


class AbstractPHS(AbstractRadialBasis):
    order = None

    def __repr__(self):
        sup = {1: "¹", 3: "³", 5: "⁵", 7: "⁷"}[self.order]
        return f"Polyharmonic spline (r{sup})\n└─Polynomial augmentation: degree {self.poly_deg}"


class PHS1(AbstractPHS):
    kind, order = abi.RBFX_PHS1, 1


class PHS3(AbstractPHS):
    kind, order = abi.RBFX_PHS3, 3


class PHS5(AbstractPHS):
    kind, order = abi.RBFX_PHS5, 5


class PHS7(AbstractPHS):
    kind, order = abi.RBFX_PHS7, 7


def PHS(n=3, poly_deg=2):
    check_poly_deg(poly_deg)
    if n % 2 == 0 or n > 7 or n < 1:
        raise ArgumentError(f"n must be 1, 3, 5, or 7 (got n={n}). Use PHS(3) for cubic or PHS(5) for quintic.")
    return {1: PHS1, 3: PHS3, 5: PHS5, 7: PHS7}[n](poly_deg)


class IMQ(AbstractRadialBasis):
    kind = abi.RBFX_IMQ

    def __init__(self, eps=1.0, poly_deg=2):
        if not eps > 0:
            raise ArgumentError(f"Shape parameter ε must be > 0 (got ε={eps}). Typical values range from 0.1 to 10.0.")
        self.eps = float(eps)
        AbstractRadialBasis.__init__(self, poly_deg)

    def __repr__(self):
        return (f"Inverse Multiquadrics, 1/sqrt((r*ε)²+1)\n├─Shape factor: ε = {self.eps}\n"
                f"└─Polynomial augmentation: degree {self.poly_deg}")


class Gaussian(AbstractRadialBasis):
    kind = abi.RBFX_GAUSSIAN

    def __init__(self, eps=1.0, poly_deg=2):
        if not eps > 0:
            raise ArgumentError(f"Shape parameter ε must be > 0 (got ε={eps}). Typical values range from 0.1 to 10.0.")
        self.eps = float(eps)
        AbstractRadialBasis.__init__(self, poly_deg)

    def __repr__(self):
        return (f"Gaussian, exp(-(ε*r)²)\n├─Shape factor: ε = {self.eps}\n"
                f"└─Polynomial augmentation: degree {self.poly_deg}")


class MonomialBasis:
    """Monomials of total degree <= deg in `dim` variables."""

    def __init__(self, dim, deg):
        if deg < 0:
            raise ArgumentError("Monomial basis must have non-negative degree")
        self.dim, self.deg = int(dim), int(deg)

    @property
    def n_terms(self):
        return comb(self.dim + self.deg, self.dim)

    def __repr__(self):
        return f"MonomialBasis of degree {self.deg} in {self.dim} dimensions"


def n_monomials(dim, deg):
    return 0 if deg < 0 else comb(dim + deg, dim)


# ---------------------------------------------------------------- boundary conditions
class BoundaryCondition:
    """Bu = α u + β ∂ₙu (types.jl:18-26)."""

    __slots__ = ("alpha", "beta")

    def __init__(self, alpha, beta):
        self.alpha, self.beta = float(alpha), float(beta)

    def __repr__(self):
        return f"BoundaryCondition(α={self.alpha}, β={self.beta})"

    def __eq__(self, other):
        return isinstance(other, BoundaryCondition) and (self.alpha, self.beta) == (other.alpha, other.beta)


def Dirichlet():
    return BoundaryCondition(1.0, 0.0)


def Neumann():
    return BoundaryCondition(0.0, 1.0)


def Robin(alpha, beta):
    return BoundaryCondition(alpha, beta)


def Internal():
    return BoundaryCondition(0.0, 0.0)


def is_dirichlet(bc):
    return bc.alpha == 1.0 and bc.beta == 0.0


def is_neumann(bc):
    return bc.alpha == 0.0 and bc.beta == 1.0


def is_robin(bc):
    return bc.alpha != 0.0 and bc.beta != 0.0


def is_internal(bc):
    return bc.alpha == 0.0 and bc.beta == 0.0
