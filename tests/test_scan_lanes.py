"""CPU checks of the lane logic of the wave-level segmented scans in
fixedeffects.jl_amd/csrc/linmap.h (seg_scan): the K = 1 form built from DPP
row shifts and row broadcasts, and the K = 32 form built from a swap of the
wave's 32-lane halves, emulated lane by lane with the documented move
semantics and compared with a direct segmented inclusive scan.  (The
hardware moves themselves are exercised by every D'u parity test on the
GPU.)"""
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
