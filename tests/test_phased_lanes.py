"""Lane-by-lane CPU emulation of the phased few-level accumulators of the
fused pass (linmap.h phased_add) for a wave of 64 lanes, K columns
(lane = psub*K + c, P = 64/K positions per wave step), one accumulator per
(level, column) that the P lanes of a column add to in P ordered phases.

It covers a phased read-modify-write of compensated float pairs (Knuth
two-sum, the F32x2 of linmap.h), the form an FP32 variant used before it
failed on the GPU:
* executed phase by phase, every addition lands, and the compensated sums
  match an exact sum far below one float rounding;
* with the phases merged into one instruction (what the compiler did to the
  FP64 atomic form at K = 32), lanes of one column that meet at a level read
  the same old value and only one write survives: additions are lost;
* read back as doubles (the hand-off the FP64 form uses when it combines the
  waves' accumulators), a pair of floats left in place is not its value.
The level sums are those of the reference's gather! (src/SolverCPU.jl:21-27).
"""
import numpy as np

f32 = np.float32


def two_sum_pair(a, b):
    """F32x2 operator+ of linmap.h, in float32 arithmetic."""
    ah, al = a
    bh, bl = b
    s = f32(ah + bh)
    bp = f32(s - ah)
    e = f32(f32(ah - f32(s - bp)) + f32(bh - bp))
    lo = f32(f32(al + bl) + e)
    h = f32(s + lo)
    return (h, f32(lo - f32(h - s)))


def emulate(K, slots, x, merged=False):
    """slots[step, lane], x[step, lane] (float32): every wave step adds lane
    values into acc[slot, c]; returns the accumulators as (hi, lo)."""
    P = 64 // K
    nlev = int(slots.max()) + 1
    acc = [[(f32(0), f32(0)) for _ in range(K)] for _ in range(nlev)]
    for st in range(slots.shape[0]):
        if merged:
            # one instruction: every lane reads first, then every lane writes
            # (the last writer of an address wins)
            reads = [acc[slots[st, l]][l % K] for l in range(64)]
            for l in range(64):
                acc[slots[st, l]][l % K] = two_sum_pair(reads[l], (x[st, l], f32(0)))
        else:
            for ph in range(P):
                for l in range(64):
                    if l // K == ph:
                        k, c = slots[st, l], l % K
                        acc[k][c] = two_sum_pair(acc[k][c], (x[st, l], f32(0)))
    return acc, nlev


def _data(K, steps=200, nlev=20, seed=3):
    rng = np.random.default_rng(seed)
    slots = rng.integers(0, nlev, (steps, 64))
    x = (rng.standard_normal((steps, 64)) * np.exp(rng.uniform(-8, 8, (steps, 64)))).astype(np.float32)
    return slots, x


def _exact(K, slots, x, nlev):
    out = np.zeros((nlev, K))
    for st in range(slots.shape[0]):
        for l in range(64):
            out[slots[st, l], l % K] += float(x[st, l])  # float64 sums of float32 values
    return out


def test_phased_compensated_sums_are_accurate():
    for K in (16, 32):
        slots, x = _data(K)
        acc, nlev = emulate(K, slots, x)
        got = np.array([[float(h) + float(lo) for (h, lo) in row] for row in acc])
        want = _exact(K, slots, x, nlev)
        scale = np.abs(x).max() * slots.shape[0]
        assert np.abs(got - want).max() <= 1e-12 * scale, np.abs(got - want).max() / scale


def test_merged_phases_lose_updates():
    for K in (16, 32):
        slots, x = _data(K)
        acc, nlev = emulate(K, slots, x, merged=True)
        got = np.array([[float(h) + float(lo) for (h, lo) in row] for row in acc])
        want = _exact(K, slots, x, nlev)
        scale = np.abs(x).max() * slots.shape[0]
        assert np.abs(got - want).max() > 1e-3 * scale


def test_pair_read_as_double_is_not_its_value():
    slots, x = _data(32)
    acc, nlev = emulate(32, slots, x)
    pairs = np.array([[(h, lo) for (h, lo) in row] for row in acc], dtype=np.float32)
    as_double = pairs.reshape(-1, 2).view(np.float64).reshape(nlev, 32)
    want = _exact(32, slots, x, nlev)
    # far from the sums (tiny or huge magnitudes, sometimes not finite)
    bad = ~np.isfinite(as_double) | (np.abs(as_double - want) > 1e-3 * np.abs(want))
    assert bad.mean() > 0.9
