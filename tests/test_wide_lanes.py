"""Lane-by-lane CPU emulation of the fused pass's wide lane map
(fixedeffects.jl_amd/csrc/linmap.h, k_fused with C > 1): every lane holds C
adjacent columns of one position (lane = psub*KL + cg, columns cg*C .. cg*C+C-1,
KL = K/C lanes and P = 64/KL positions per wave step).  Emulated for FP64
batches of 16 and 32 columns with C = 2 and 4:

* seg_step_c: the head-bit logic once per position, then per column the
  carry-in, the segmented scan over the step's positions (shifts by multiples
  of KL lanes), the emission of a segment's sum at its last position into
  sums[segment*K + cg*C + j], and the carry of column cg*C + j's last
  position into the next step;
* phased_add_c: in each phase the active lanes hold distinct columns (no two
  lanes of one LDS add meet), and every (position, column) is added once;
* the norm partials: per lane and column, summed over the lanes of a column
  group (butterfly over multiples of KL), then thread cg writes column cg*C + j.

Checked against direct sums over (segment, column) -- the D'u level sums of
the reference's gather! (src/SolverCPU.jl:21-27) -- with planted mistakes in
the lane map that must fail."""
import numpy as np
import pytest

W = 64


def _heads_words(keys, n):
    nwords = n // 64 + 2
    h = np.ones(nwords * 64, bool)
    h[1:n] = keys[1:n] != keys[:n - 1]
    words = [0] * nwords
    for p in np.nonzero(h)[0]:
        words[p >> 6] |= 1 << int(p & 63)
    return words


def _shfl_up(v, d):
    """__shfl_up(v, d): lane l reads lane l - d (lanes below d keep theirs)."""
    return np.array([v[l - d] if l >= d else v[l] for l in range(W)])


def emulate_tile(un, seg, words, start, end, K, C, lane_col=None, scan_stride=None):
    """One wave's walk over positions [start, end) (start a multiple of 64,
    every segment ends inside the tile, as pieces do).  un[pos, col] are the
    new u values; seg[pos] is the position's segment ordinal.  Returns the
    emitted sums {(segment, column): value}."""
    KL = K // C
    P = W // KL
    lane = np.arange(W)
    psub, cg = lane // KL, lane % KL
    c0 = cg * C if lane_col is None else lane_col(cg)
    stride = KL if scan_stride is None else scan_stride
    out = {}
    carry = np.zeros((W, C))
    for base in range(start, end, P):
        w = words[base >> 6]
        sh = base & 63
        hb = (w >> sh) & ((1 << P) - 1)
        nb = (w >> (sh + P)) & 1 if sh + P < 64 else words[(base >> 6) + 1] & 1
        pos = base + psub
        valid = pos < end
        x = np.zeros((W, C))
        for l in range(W):
            if valid[l]:
                x[l] = un[pos[l], c0[l]:c0[l] + C]
        rs = np.array([max([q for q in range(psub[l] + 1) if (hb >> q) & 1], default=0) for l in range(W)])
        for l in range(W):
            if psub[l] == 0 and not (hb & 1):
                x[l] += carry[l]
        for j in range(C):
            v = x[:, j].copy()
            d = 1
            while d < P:
                up = _shfl_up(v, d * stride)
                v = v + np.where(psub - d >= rs, up, 0.0)
                d <<= 1
            x[:, j] = v
        for l in range(W):
            ps = psub[l]
            isend = nb if ps == P - 1 else (hb >> (ps + 1)) & 1
            if valid[l] and isend:
                for j in range(C):
                    key = (int(seg[pos[l]]), int(c0[l] + j))
                    assert key not in out
                    out[key] = x[l, j]
        is_open = nb == 0
        for l in range(W):
            src = (P - 1) * KL + cg[l]
            carry[l] = x[src] if is_open else 0.0
    return out


def phased_adds(slots, xs, K, C):
    """phased_add_c over one wave step: acc[slot*K + c0 + j] += x[lane, j] in
    P phases; asserts that the lanes of one add never share an address."""
    KL = K // C
    P = W // KL
    lane = np.arange(W)
    psub, cg = lane // KL, lane % KL
    acc = {}
    for ph in range(P):
        act = lane[psub == ph]
        for j in range(C):
            addrs = [slots[l] * K + cg[l] * C + j for l in act]
            assert len(set(addrs)) == len(addrs)
            for l, a in zip(act, addrs):
                acc[a] = acc.get(a, 0.0) + xs[l, j]
    return acc


def block_partials(sq, K, C, nwaves=4):
    """sq[wave, lane, j] -> partials[column] as block_column_sum<KL> per j."""
    KL = K // C
    out = np.zeros(K)
    for j in range(C):
        red = np.zeros((nwaves, KL))
        for w in range(nwaves):
            v = sq[w, :, j].copy()
            d = KL
            while d < W:
                v = v + np.array([v[l ^ d] for l in range(W)])
                d <<= 1
            red[w] = v[:KL]
        for t in range(KL):
            out[t * C + j] = red[:, t].sum()
    return out


def _panel(rng, n, K):
    sizes = rng.geometric(1 / rng.choice([2, 5, 20]), size=n)
    keys = np.repeat(np.arange(len(sizes)), sizes)[:n]
    R = 256
    # pieces: runs cut at every tile start (every tile starts a segment)
    cut = np.zeros(n, bool)
    cut[::R] = True
    cut[1:] |= keys[1:] != keys[:-1]
    cut[0] = True
    seg = np.cumsum(cut) - 1
    un = rng.standard_normal((n, K))
    return seg, un, R


@pytest.mark.parametrize("K,C", [(16, 2), (16, 4), (32, 2), (32, 4)])
def test_wide_segment_sums(K, C):
    rng = np.random.default_rng(K * 10 + C)
    n = 1024 + int(rng.integers(0, 200))
    seg, un, R = _panel(rng, n, K)
    words = _heads_words(seg, n)
    got = {}
    for start in range(0, n, R):
        got.update(emulate_tile(un, seg, words, start, min(start + R, n), K, C))
    want = np.zeros((seg.max() + 1, K))
    np.add.at(want, seg, un)
    assert len(got) == want.size
    for (s, c), v in got.items():
        assert abs(v - want[s, c]) <= 1e-12 * (1 + abs(want[s, c]))


@pytest.mark.parametrize("K,C", [(16, 4), (32, 4)])
def test_planted_lane_map_mistakes_fail(K, C):
    rng = np.random.default_rng(5)
    n = 700
    seg, un, R = _panel(rng, n, K)
    words = _heads_words(seg, n)
    want = np.zeros((seg.max() + 1, K))
    np.add.at(want, seg, un)

    def wrong(**kw):
        got = {}
        try:
            for start in range(0, n, R):
                got.update(emulate_tile(un, seg, words, start, min(start + R, n), K, C, **kw))
        except (AssertionError, ValueError, IndexError):
            return True
        return len(got) != want.size or any(abs(v - want[s, c]) > 1e-9 for (s, c), v in got.items())

    assert wrong(lane_col=lambda cg: cg)  # column group taken as the first column
    assert wrong(scan_stride=1)  # scan shifts by lanes instead of positions


@pytest.mark.parametrize("K,C", [(16, 2), (16, 4), (32, 2), (32, 4)])
def test_wide_phased_adds_and_partials(K, C):
    rng = np.random.default_rng(7 + K + C)
    nslot = 20
    KL = K // C
    P = W // KL
    total = np.zeros(nslot * K)
    want = np.zeros(nslot * K)
    for step in range(30):
        slots = rng.integers(0, nslot, W)
        xs = rng.standard_normal((W, C))
        for a, v in phased_adds(slots, xs, K, C).items():
            total[a] += v
        for l in range(W):
            for j in range(C):
                want[slots[l] * K + (l % KL) * C + j] += xs[l, j]
    np.testing.assert_allclose(total, want, rtol=1e-12, atol=1e-12)
    # norm partials: column cg*C + j of lane (psub, cg)
    sq = rng.random((4, W, C))
    got = block_partials(sq, K, C)
    direct = np.zeros(K)
    for w in range(4):
        for l in range(W):
            for j in range(C):
                direct[(l % KL) * C + j] += sq[w, l, j]
    np.testing.assert_allclose(got, direct, rtol=1e-12)
