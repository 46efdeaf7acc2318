"""fixedeffects_rocm -- FixedEffects.jl's residualization path on AMD MI355X.

The solve runs in libfehip.so (hand-written gfx950 HIP kernels behind a C-ABI,
include/fehip.h).  This package mirrors the reference's public API for the
``:rocm`` backend; see ``api.py`` for the citations.
"""
from ._lib import FehipError, LIB_PATH  # noqa: F401
from .api import (  # noqa: F401
    AbstractFixedEffectSolver,
    DimensionMismatch,
    FixedEffect,
    FixedEffectSolverROCm,
    components,
    full,
    normalize,
    solve_coefficients,
    solve_residuals,
)
