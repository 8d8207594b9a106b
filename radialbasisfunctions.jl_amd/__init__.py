"""MI355X-native RBF-FD weight engine with the operator API of RadialBasisFunctions.jl.

The weight build (k-NN, per-stencil assemble-and-solve) and the sparse apply run
in librbfx.so, a C-ABI library of hand-written gfx950 HIP kernels (include/rbfx.h).
This package is the host-side mirror of the reference's operator layer.
"""
from . import _abi
from . import _lib
from . import learned
from .basis import (PHS, PHS1, PHS3, PHS5, PHS7, IMQ, Gaussian, MonomialBasis, AbstractRadialBasis, AbstractPHS,
                    BoundaryCondition, Dirichlet, Neumann, Robin, Internal, is_dirichlet, is_neumann, is_robin,
                    is_internal, check_poly_deg, n_monomials)
from .errors import ArgumentError, DimensionMismatch, SingularException
from .operators import (AbstractOperator, RadialBasisOperator, Laplacian, Partial, MixedPartial, Jacobian, Hessian,
                        Regrid, Identity, ScaledOperator, SumOperator, Custom, Stacked, Hermite, laplacian, partial,
                        mixed_partial, jacobian, gradient, hessian, regrid, custom, find_neighbors, autoselect_k,
                        reorder_points,
                        weights, update_weights, is_cache_valid, print_op)
from .learned import LearnedModel
from . import functors
from .notation import operator

__all__ = [n for n in dir() if not n.startswith("_")]
