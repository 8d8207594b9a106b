"""Exception types matching the ones the reference raises on this path."""
import numpy as np


class ArgumentError(ValueError):
    """Julia ArgumentError (e.g. src/solve/api.jl:382-397, src/utils.jl:57-66)."""


class DimensionMismatch(ValueError):
    """Julia DimensionMismatch (src/solve/api.jl:399-419, src/solve/types.jl:82-96)."""


class SingularException(np.linalg.LinAlgError):
    """LinearAlgebra.SingularException raised by bunchkaufman! (src/solve/assembly.jl:220-222)."""
