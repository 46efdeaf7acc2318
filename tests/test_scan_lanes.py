"""CPU checks of the lane logic of the wave-level segmented scans in
fixedeffects.jl_amd/csrc/linmap.h (seg_scan): the K = 1 form built from DPP
row shifts and row broadcasts, and the K = 32 form built from a swap of the
wave's 32-lane halves, emulated lane by lane with the documented move
semantics and compared with a direct segmented inclusive scan.  (The
hardware moves themselves are exercised by every D'u parity test on the
GPU.)  Also emulated: the FE s level from the head words (k_fused level_s),
and the CSR tile walk (dtx_seg_body with clipped head bits, seg_step, and the
fix-up's combination of tile partials) for tiles cut at segment heads, at
arbitrary positions and at fixed sizes."""
import numpy as np


def reference_scan(x, heads, K):
    """Inclusive scan of each column over the P = 64/K positions of a wave
    step, restarting at heads (lane = psub*K + c)."""
    P = 64 // K
    out = np.zeros(64)
    for c in range(K):
        run = 0.0
        for ps in range(P):
            lane = ps * K + c
            run = x[lane] if heads[ps] else run + x[lane]
            out[lane] = run
    return out


def latest_head(heads, P):
    rs = np.zeros(P, int)
    last = 0
    for ps in range(P):
        if heads[ps]:
            last = ps
        rs[ps] = last
    return rs


def row_shr(v, n):  # row_shr:n, lanes without a source in their row read 0
    return np.array([v[l - n] if l % 16 >= n else 0.0 for l in range(64)])


def row_bcast15(v):  # rows 1 and 3 read lanes 15 and 47 (row mask 0xA)
    return np.array([v[(l // 16) * 16 - 1] if (l // 16) in (1, 3) else 0.0 for l in range(64)])


def row_bcast31(v):  # rows 2 and 3 read lane 31 (row mask 0xC)
    return np.array([v[31] if l >= 32 else 0.0 for l in range(64)])


def half_swap(v):  # lane l reads lane (l + 32) mod 64
    return np.array([v[(l + 32) % 64] for l in range(64)])


def scan_k1(x, heads):
    lane = np.arange(64)
    rs = latest_head(heads, 64)
    v = x.copy()
    for d in (1, 2, 4, 8):
        up = row_shr(v, d)
        v = v + np.where(((lane & 15) >= d) & (lane - d >= rs), up, 0.0)
    up = row_bcast15(v)
    v = v + np.where(((lane & 16) != 0) & (rs < (lane & ~15)), up, 0.0)
    up = row_bcast31(v)
    v = v + np.where((lane >= 32) & (rs < 32), up, 0.0)
    return v


def scan_k32(x, heads):
    psub = np.arange(64) // 32
    rs = latest_head(heads, 2)[psub]
    up = half_swap(x)
    return x + np.where((psub == 1) & (rs == 0), up, 0.0)


def test_dpp_scan_k1_matches_segmented_scan():
    rng = np.random.default_rng(0)
    for trial in range(3000):
        x = rng.standard_normal(64)
        heads = rng.random(64) < rng.choice([0.0, 0.02, 0.1, 0.5, 1.0])
        np.testing.assert_allclose(scan_k1(x, heads), reference_scan(x, heads, 1), rtol=1e-12, atol=1e-12)


def test_half_swap_scan_k32_matches_segmented_scan():
    rng = np.random.default_rng(1)
    for trial in range(500):
        x = rng.standard_normal(64)
        heads = rng.random(2) < 0.5
        np.testing.assert_allclose(scan_k32(x, heads), reference_scan(x, heads, 32), rtol=1e-12, atol=1e-12)
    # the carry into the next step is column c's last position (lane 32 + c)
    x = rng.standard_normal(64)
    last = np.where(np.arange(64) < 32, half_swap(x), x)
    np.testing.assert_array_equal(last[:32], x[32:])
    np.testing.assert_array_equal(last[32:], x[32:])


# -- FE s level from the head words (k_fused level_s) ------------------------------

def _heads_and_prefix(keys, n):
    """Head bits (bit j of word j//64: j == 0, a key change, or j >= n) and the
    per-word count of heads before the word (positions < n), as csr_build.h
    builds them."""
    nwords = n // 64 + 2
    pos = np.arange(nwords * 64)
    h = np.ones(nwords * 64, bool)
    h[1:n] = keys[1:n] != keys[:n - 1]
    real = h.copy()
    real[n:] = False
    words = np.zeros(nwords, np.uint64)
    for b in range(64):
        words |= (h[pos % 64 == b].astype(np.uint64) << np.uint64(b))
    cnt = real.reshape(nwords, 64).sum(1)
    wprefix = np.concatenate([[0], np.cumsum(cnt)])[:nwords].astype(np.uint32)
    return words, wprefix


def _level_s(words, wprefix, seg_level, nsegs, start, end, lb, q, psub, P, NW):
    """k_fused level_s / load_words_wp, lane by lane (uint32 wrap-around and
    clamping as in the kernel)."""
    length = end - start
    b0 = start + lb
    jw = (q * P) // 64
    wi = (b0 >> 6) + jw
    ok = lb < length and wi <= ((end - 1) >> 6) + 1
    hw = int(words[wi]) if ok else 0
    wp = int(wprefix[wi]) if ok else 0
    bit = (lb + q * P + psub) & 63
    le = (1 << (bit + 1)) - 1
    ordv = (wp + bin(hw & le).count("1") - 1) & 0xFFFFFFFF
    ordv = ordv if ordv < nsegs else nsegs - 1
    return int(seg_level[ordv]) if seg_level is not None else ordv


def test_level_from_head_words():
    rng = np.random.default_rng(3)
    for trial in range(6):
        n = int(rng.integers(300, 3000))
        nlev = int(rng.integers(5, n // 2))
        keys = np.sort(rng.integers(0, nlev, n))
        if trial % 2:  # drop some levels entirely (seg_level not the identity)
            keys = np.unique(keys, return_inverse=True)[1] * 2
        words, wprefix = _heads_and_prefix(keys, n)
        present = np.unique(keys)
        identity = len(present) == keys.max() + 1
        seg_level = None if identity else present
        nsegs = len(present)
        for K, U in ((1, 2), (1, 4), (2, 2), (4, 2), (8, 2), (16, 2), (32, 2)):
            P = 64 // K
            UP = U * P
            NW = max(UP // 64 + 1, 2)
            R = 256 * int(rng.integers(1, 4))
            for start in range(0, n, R):
                end = min(start + R, n)
                for lb in range(0, end - start, UP):
                    for q in range(U):
                        for psub in range(P):
                            posn = start + lb + q * P + psub
                            if posn >= end:
                                continue
                            got = _level_s(words, wprefix, seg_level, nsegs, start, end, lb, q, psub, P, NW)
                            assert got == keys[posn], (trial, K, U, start, lb, q, psub)


# -- CSR tile walk (linmap.h dtx_seg_body + seg_step + the fix-up's kind-0 items) --

def _clip(w, wi, start, end):
    """The head bits of positions [start, end] of word wi (dtx_seg_body clip)."""
    lo, hi = int(start) - (int(wi) << 6), int(end) - (int(wi) << 6)
    if hi < 0 or lo > 63:
        return 0
    m = ((1 << 64) - 1) << lo if lo > 0 else (1 << 64) - 1
    if hi < 63:
        m &= (2 << hi) - 1
    return w & m & ((1 << 64) - 1)


def _seg_step(w, wnext, sh, valid, x, cur, P):
    """seg_step for one column, lane by lane: returns (scanned x, lanes that
    emit, lanes that write tileA)."""
    hb = w if P == 64 else (w >> sh) & ((1 << P) - 1)
    nb = (w >> (sh + P)) & 1 if sh + P < 64 else wnext & 1
    x = x.copy()
    if not (hb & 1):
        x[0] += cur["carry"]
    out = np.zeros(P)
    run = 0.0
    for ps in range(P):
        run = x[ps] if (hb >> ps) & 1 else run + x[ps]
        out[ps] = run
    em, ta = [], []
    for ps in range(P):
        hle = hb & ((2 << ps) - 1)
        isend = nb if ps == P - 1 else (hb >> (ps + 1)) & 1
        if valid[ps] and isend:
            (ta if (cur["cont"] and hle == 0) else em).append(ps)
    cur["open"] = nb == 0
    cur["carry"] = out[P - 1] if cur["open"] else 0.0
    cur["cont"] = cur["cont"] and hb == 0
    return out, em, ta


def _walk(words, val, level_of, tiles, P):
    """Every tile walked as the kernel does (from the word holding its start,
    P positions a step); returns the directly written level sums and the
    tileA / tileB partials."""
    direct, tileA, tileB = {}, {}, {}
    for t, (start, end) in enumerate(tiles):
        start, end = int(start), int(end)
        cur = {"carry": 0.0, "cont": True, "open": False}
        for base in range(start & ~63, end, P):
            wi = base >> 6
            w, wn = _clip(int(words[wi]), wi, start, end), _clip(int(words[wi + 1]), wi + 1, start, end)
            pos = base + np.arange(P)
            valid = (pos >= start) & (pos < end)
            x = np.where(valid, val[np.clip(pos, 0, len(val) - 1)], 0.0)
            out, em, ta = _seg_step(w, wn, base & 63, valid, x, cur, P)
            for ps in em:
                k = level_of[pos[ps]]
                assert k not in direct
                direct[k] = out[ps]
            for ps in ta:
                tileA[t] = out[ps]
        if cur["open"]:
            (tileA if cur["cont"] else tileB)[t] = cur["carry"]
    return direct, tileA, tileB


def _tiles_at_heads(off, n, tp):
    """CSR tile starts as the plan cuts them: at the last head at or before
    each target, or at the target inside a segment longer than a tile."""
    q, starts = 0, []
    while q < n:
        starts.append(q)
        tgt = q + tp
        if tgt >= n:
            break
        head = off[np.searchsorted(off, tgt, side="right") - 1]
        q = head if head > q else tgt
    starts.append(n)
    return starts


def test_csr_tile_walk_at_heads_and_inside_segments():
    # segment sums from the walk + the fix-up's combination equal the direct
    # sums, for tiles cut at heads (the plan's cut), tiles cut anywhere, and
    # the old fixed tiles; long segments span several tiles
    rng = np.random.default_rng(11)
    for trial in range(8):
        n = int(rng.integers(500, 4000))
        sizes = rng.geometric(1 / rng.choice([3, 20, 150]), size=n)
        sizes[rng.integers(0, len(sizes), 2)] += int(rng.integers(300, 900))  # long segments
        keys = np.repeat(np.arange(len(sizes)), sizes)[:n]
        nlev = keys.max() + 1
        off = np.concatenate([[0], np.cumsum(np.bincount(keys, minlength=nlev))])
        words, _ = _heads_and_prefix(keys, n)
        val = rng.standard_normal(n)
        want = np.bincount(keys, weights=val, minlength=nlev)
        tp = 256 * int(rng.integers(1, 3))
        cuts = {
            "heads": _tiles_at_heads(off, n, tp),
            "anywhere": sorted(set([0, n] + list(rng.choice(np.arange(1, n), size=n // 97, replace=False)))),
            "fixed": list(range(0, n, tp)) + [n],
        }
        for name, st in cuts.items():
            tiles = list(zip(st[:-1], st[1:]))
            tstart = np.array(st)
            for K in (1, 4, 32):
                P = 64 // K
                direct, tileA, tileB = _walk(words, val, keys, tiles, P)
                crossing = 0
                for k in range(nlev):
                    a, b = off[k], off[k + 1]
                    tf = np.searchsorted(tstart, a, side="right") - 1
                    tl = np.searchsorted(tstart, b - 1, side="right") - 1
                    if tf == tl:
                        got = direct[k]
                    else:
                        crossing += 1
                        got = tileB[tf] + sum(tileA[q] for q in range(tf + 1, tl + 1))
                    assert abs(got - want[k]) <= 1e-9 * (1 + abs(want[k])), (trial, name, K, k)
                if name == "heads":  # only segments longer than a tile cross
                    assert crossing <= sum((off[k + 1] - off[k]) > tp for k in range(nlev)), (trial, crossing)
