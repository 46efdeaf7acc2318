"""CPU emulation of the piece form of D'u (the plain LSMR iteration of a
factored handle whose only CSR FE is the secondary sort key t).

In the internal order (observations sorted by FE s level, then FE t level), a
piece is a run of one (s, t) level pair cut at every fused-pass tile start
(multiples of R).  The fused pass writes each piece's sum; the piece launch
then forms FE s's level sums over its contiguous pieces and FE t's over the
pieces ordered stably by t level, both in tiles cut at level heads whose
crossing levels the fix-up completes.  This mirrors k_piece_heads,
k_piece_levels, build_pieces / cut_at_heads and the D'u reference gather!
(src/SolverCPU.jl:21-27), whose per-level sums are the target: checked
exactly on integer-valued data (every partial sum is exact), with planted
mistakes that must fail.
"""
import numpy as np
import pytest

from fixedeffects_rocm import synth


def _internal_order(s, t):
    return np.lexsort((np.arange(len(s)), t, s))  # by s, then t, then index (stable)


def _pieces(s, t, R, tile_heads=True, ord_shift=0):
    N = len(s)
    j = np.arange(N)
    head = np.ones(N, bool)
    head[1:] = (s[1:] != s[:-1]) | (t[1:] != t[:-1])
    if tile_heads:
        head |= (j % R) == 0
    ordv = np.cumsum(head) - 1 + ord_shift
    np_ = int(head.sum())
    piece_s = np.zeros(np_, np.int64)
    piece_t = np.zeros(np_, np.int64)
    hs = np.flatnonzero(head)
    piece_s[np.clip(ordv[hs], 0, np_ - 1)] = s[hs]
    piece_t[np.clip(ordv[hs], 0, np_ - 1)] = t[hs]
    return head, np.clip(ordv, 0, np_ - 1), piece_s, piece_t


def _cut_at_heads(off, n, tp):
    ts, q = [], 0
    while q < n:
        ts.append(q)
        tgt = q + tp
        if tgt >= n:
            break
        head = off[np.searchsorted(off, tgt, side="right") - 1]
        q = head if head > q else tgt
    ts.append(n)
    return np.array(ts)


def _level_sums_in_tiles(vals, keys, nlev, tp):
    """Segmented sums of vals (sorted by keys) in tiles cut at heads: each
    tile sums the segments inside it, a crossing segment leaves partials
    that the fix-up adds in tile order."""
    off = np.searchsorted(keys, np.arange(nlev + 1))
    ts = _cut_at_heads(off, len(keys), tp)
    out = np.zeros(nlev)
    for a, b in zip(ts[:-1], ts[1:]):
        np.add.at(out, keys[a:b], vals[a:b])  # within a tile: per-segment partials, combined in order
    return out, ts, off


def _emulate(p, R, tp, **kw):
    s0 = p.refs[0] - 1
    t0 = p.refs[1] - 1
    order = _internal_order(s0, t0)
    s, t = s0[order], t0[order]
    rng = np.random.default_rng(4)
    x = rng.integers(-1000, 1000, len(s)).astype(np.float64)  # exact sums
    head, ordv, piece_s, piece_t = _pieces(s, t, R, **kw)
    psum = np.bincount(ordv, weights=x, minlength=len(piece_s))
    sum_s, _, _ = _level_sums_in_tiles(psum, piece_s, p.n[0], tp)
    perm_t = np.argsort(piece_t, kind="stable")
    sum_t, _, _ = _level_sums_in_tiles(psum[perm_t], piece_t[perm_t], p.n[1], tp)
    want_s = np.bincount(s, weights=x, minlength=p.n[0])
    want_t = np.bincount(t, weights=x, minlength=p.n[1])
    return (sum_s, want_s), (sum_t, want_t), head, ordv, len(piece_s), s


@pytest.mark.parametrize("R,tp", [(256, 256), (512, 256), (1024, 2048)])
def test_piece_sums_equal_level_sums(R, tp):
    p = synth.config("worker_firm_1e7", scale=0.005)  # N = 5e4, Zipf workers, mobile firms
    (gs, ws), (gt, wt), head, ordv, npc, s = _emulate(p, R, tp)
    np.testing.assert_array_equal(gs, ws)
    np.testing.assert_array_equal(gt, wt)
    # no piece crosses a fused tile: every tile start is a head, and the
    # ordinals of a tile's positions stay inside the tile's own pieces
    N = len(s)
    starts = np.arange(0, N, R)
    assert head[starts].all()
    assert npc <= N
    for a in starts:
        b = min(a + R, N)
        assert ordv[a] == (ordv[a - 1] + 1 if a else 0)
        assert ordv[b - 1] < (ordv[b] if b < N else npc)


def test_planted_mistakes_fail():
    p = synth.config("worker_firm_1e7", scale=0.005)
    # an ordinal off by one: the sums no longer match
    (gs, ws), (gt, wt), *_ = _emulate(p, 256, 256, ord_shift=1)
    assert not (np.array_equal(gs, ws) and np.array_equal(gt, wt))
    # without the tile-start heads some piece crosses a fused tile
    *_, head, ordv, npc, s = _emulate(p, 256, 256, tile_heads=False)
    starts = np.arange(0, len(s), 256)
    assert not head[starts].all()


def test_pieces_fewer_than_observations_on_the_bench_shape():
    # the form is used when pieces are at most half the observations; on the
    # worker/firm panel (mobility 0.05) a worker's observations at one firm
    # form one run in the internal order
    p = synth.config("worker_firm_1e7", scale=0.02)
    s0, t0 = p.refs[0] - 1, p.refs[1] - 1
    order = _internal_order(s0, t0)
    _, _, piece_s, _ = _pieces(s0[order], t0[order], 3328)
    assert 2 * len(piece_s) <= p.N
