"""CPU restatement of FixedEffects.jl v3.1.0's residualization path.

TEST INFRASTRUCTURE ONLY.  This package is the oracle the HIP library is
checked against (tests/, ``__graft_entry__.smoke``) and the CPU baseline that
``bench.py`` times beside it.  The product path (``fixedeffects_rocm`` and
``libfehip.so``) never imports it.

Pinning status of each function is recorded in ``tests/test_oracle.py``:

* ``bucketize``            -- pinned: stable argsort / bincount / cumsum.
* ``scale``, ``cache``     -- pinned: closed form via bincount; Jacobi
  property diag(A'A) = 1 on observed levels; issue #97 zero case.
* ``gather``, ``scatter``  -- pinned: explicit scipy.sparse matrix products
  and the adjoint identity <Ax, y> = <x, A'y>.
* ``lsmr``                 -- pinned: iterates against scipy's LSMR (the same
  Fong-Saunders algorithm) on the explicit matrix, and iteration counts from a
  zero start where both use the same normb.
* ``solve_residuals``      -- pinned: the reference's golden vector
  (test/solve.jl:3-16), dense least squares, single-FE closed form.
* ``solve_coefficients``, ``normalize``, ``components``, ``full`` -- pinned:
  golden check x - coefs[1] - coefs[2] = r_ols (test/solve.jl:9-12),
  scipy.sparse.csgraph connected components, zero component means.
"""
from .oracle import (  # noqa: F401
    OracleSolver,
    bucketize,
    components,
    full,
    lib,
    normalize,
    solve_coefficients,
    solve_residuals,
)
