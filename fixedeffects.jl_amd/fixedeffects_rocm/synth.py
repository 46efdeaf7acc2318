"""Seeded synthetic panels of the workloads the solver is measured on.

* ``uniform_two_way``  -- N obs, FE1 uniform over n1 levels, FE2 uniform over
  n2 levels, unit weights, y = a[FE1] + c[FE2] + e  (the CPU reference case:
  N=100,000, n = (10,000, 100)).
* ``worker_firm``      -- worker / firm / year panel: worker sizes follow a
  rank-size Zipf(s) law s_k = max(2, floor(C k^-s)) with C set so the sizes sum
  to N; each worker starts at a uniform firm and moves to a new uniform firm at
  each later observation with probability ``mobility``; years are uniform.
  Rows are shuffled uniformly; levels never hit are dropped and the rest
  renumbered in sorted order (dense refs, as GroupedArray produces).
  w ~ LogNormal(0, 0.5); y = a[worker] + f[firm] + t[year] + e, all N(0,1);
  extra columns x_k = 0.5 a[worker] + 0.3 f[firm] + N(0,1).
  With ``interaction_lo/hi`` the year FE is interacted with z ~ U(lo, hi).

Everything is drawn from numpy's PCG64 with the given seed.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


@dataclass
class Panel:
    refs: List[np.ndarray]             # 1-based int64 (FixedEffect.refs)
    n: List[int]                       # levels per FE
    interactions: List[Optional[np.ndarray]]
    weights: Optional[np.ndarray]      # None = UnitWeights
    Y: np.ndarray                      # N x K, column-major (Fortran order)
    info: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return len(self.refs[0])

    @property
    def K(self) -> int:
        return self.Y.shape[1]


def _dense(ids: np.ndarray):
    """0-based ids -> dense 1-based refs in sorted id order, and the level count."""
    present = np.bincount(ids) > 0
    newid = np.cumsum(present) - 1
    return newid[ids].astype(np.int64) + 1, int(present.sum())


def uniform_two_way(N: int = 100_000, n1: int = 10_000, n2: int = 100, seed: int = 0) -> Panel:
    rng = np.random.default_rng(seed)
    f1 = rng.integers(0, n1, N)
    f2 = rng.integers(0, n2, N)
    a = rng.standard_normal(n1)
    c = rng.standard_normal(n2)
    y = a[f1] + c[f2] + rng.standard_normal(N)
    r1, m1 = _dense(f1)
    r2, m2 = _dense(f2)
    return Panel([r1, r2], [m1, m2], [None, None], None, np.asfortranarray(y[:, None]),
                 {"generator": "uniform_two_way", "seed": seed, "N": N, "n": (n1, n2)})


def zipf_sizes(N: int, n_levels: int, s: float = 1.1, min_size: int = 2) -> np.ndarray:
    """Rank-size Zipf sizes max(min_size, floor(C k^-s)), k = 1..n_levels,
    with C found by bisection so the total is as close to N as possible; the
    remainder goes to the largest level."""
    if n_levels * min_size > N:
        raise ValueError("too many levels for N")
    k = np.arange(1, n_levels + 1, dtype=np.float64)
    kp = k ** (-s)

    def total(C):
        return int(np.maximum(min_size, np.floor(C * kp)).sum())

    lo, hi = 0.0, float(N)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if total(mid) <= N:
            lo = mid
        else:
            hi = mid
    sizes = np.maximum(min_size, np.floor(lo * kp)).astype(np.int64)
    sizes[0] += N - int(sizes.sum())
    return sizes


def worker_firm(N: int = 10_000_000, n_workers: int = 1_000_000, n_firms: int = 100_000, n_years: int = 20,
                mobility: float = 0.05, zipf_s: float = 1.1, K: int = 1, seed: int = 1,
                weights: bool = True, shuffle: bool = True, interaction_lo: Optional[float] = None,
                interaction_hi: Optional[float] = None) -> Panel:
    rng = np.random.default_rng(seed)
    sizes = zipf_sizes(N, n_workers, zipf_s)
    worker = np.repeat(np.arange(n_workers, dtype=np.int64), sizes)
    first = np.zeros(N, bool)
    first[np.concatenate([[0], np.cumsum(sizes)[:-1]])] = True
    move = first | (rng.random(N) < mobility)
    events = np.cumsum(move) - 1
    firm = rng.integers(0, n_firms, int(move.sum()))[events]
    year = rng.integers(0, n_years, N)
    a = rng.standard_normal(n_workers)
    f = rng.standard_normal(n_firms)
    t = rng.standard_normal(n_years)
    Y = np.empty((N, K), order="F")
    Y[:, 0] = a[worker] + f[firm] + t[year] + rng.standard_normal(N)
    for k in range(1, K):
        Y[:, k] = 0.5 * a[worker] + 0.3 * f[firm] + rng.standard_normal(N)
    w = rng.lognormal(0.0, 0.5, N) if weights else None
    z = None
    if interaction_lo is not None:
        z = rng.uniform(interaction_lo, interaction_hi, N)
    if shuffle:
        perm = rng.permutation(N)
        worker, firm, year = worker[perm], firm[perm], year[perm]
        Y = np.asfortranarray(Y[perm])
        if w is not None:
            w = w[perm]
        if z is not None:
            z = z[perm]
    r1, m1 = _dense(worker)
    r2, m2 = _dense(firm)
    r3, m3 = _dense(year)
    info = {"generator": "worker_firm", "seed": seed, "N": N, "n": (n_workers, n_firms, n_years),
            "mobility": mobility, "zipf_s": zipf_s, "max_worker_size": int(sizes.max()),
            "workers_at_2": int((sizes == 2).sum()), "shuffled": shuffle}
    return Panel([r1, r2, r3], [m1, m2, m3], [None, None, z], w, Y, info)


def uneven_panel(seed: int, n_min: float = 1e4, n_max: float = 2e6) -> Panel:
    """One panel of the uneven-batch workload: N log-uniform in [n_min, n_max],
    2 or 3 FEs with N/10 Zipf(1.1) workers and N/100 firms (+ 20 years when
    there are 3), mobility ~ U(0.01, 0.1), 1..8 columns built as above."""
    rng = np.random.default_rng(seed)
    N = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
    G = int(rng.integers(2, 4))
    mob = float(rng.uniform(0.01, 0.1))
    K = int(rng.integers(1, 9))
    p = worker_firm(N, max(2, N // 10), max(2, N // 100), 20, mob, 1.1, K, seed=seed + 7_919)
    if G == 2:
        p.refs, p.n, p.interactions = p.refs[:2], p.n[:2], p.interactions[:2]
    p.info.update({"panel_seed": seed, "G": G})
    return p


def config(name: str, scale: float = 1.0, seed: int = 1) -> Panel:
    """The named workloads; ``scale`` shrinks N and the level counts together
    (used for parity cases the CPU oracle finishes in seconds)."""
    if name == "cpu_two_way":
        return uniform_two_way(int(100_000 * scale), max(2, int(10_000 * scale)), 100, seed=seed)
    if name in ("worker_firm_1e7", "worker_firm_1e7_k32"):
        K = 32 if name.endswith("k32") else 1
        return worker_firm(int(10_000_000 * scale), max(2, int(1_000_000 * scale)), max(2, int(100_000 * scale)),
                           20, 0.05, 1.1, K, seed)
    if name == "worker_firm_1e8_k16":
        return worker_firm(int(100_000_000 * scale), max(2, int(5_000_000 * scale)), max(2, int(500_000 * scale)),
                           50, 0.02, 1.1, 16, seed, interaction_lo=0.5, interaction_hi=1.5)
    raise KeyError(name)
