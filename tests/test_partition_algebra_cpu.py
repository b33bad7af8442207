"""Index algebra the device partition relies on, checked on the CPU with
plain Python (no GPU, no library): the identities are restated here from
the definitions, independent of the kernels' code.

* Level-1 digits from a window of bases: on the main path a k-mer's key is
  (OR_j code3(c_j) << 2(k-1-j)) masked to 2k bits (the reference's packing,
  src/kmer_counting.cpp:2973-3017, with the 3-bit code's third bit ORed into
  the next-older slot).  Its top d <= 13 bits depend only on the window's
  first seven bases, and equal bits [14-d, 14) of a 15-bit register rolled
  as ((v << 2) | code3) & 0x7FFF, read at the window's seventh base.
* Level-2 segments of the folded histogram: with the level-1 output in
  digit-major order, the keys that a run of level-1 blocks wrote for one
  level-1 bucket are contiguous; scanning per-segment counts in (bucket,
  sub-digit, segment) order and giving each segment's tiles running bases
  places every key exactly once inside its (bucket, sub-digit) range, also
  when the bases are read from the transposed (segment, sub-digit) copy."""
from __future__ import annotations

import random

import numpy as np


def code3(b: int) -> int:
    return (b >> 1) & 7


def test_window_digit_equals_key_top_bits():
    rng = random.Random(5)
    alpha = b"ACGTacgtNnRYKMSWnx-."
    for _ in range(300):
        k = rng.choice([7, 8, 9, 12, 16, 21, 31, 32, 33, 48, 63, 64])
        n = rng.randint(0, 160)
        read = bytes(rng.choice(alpha if rng.random() < 0.3 else b"ACGT") for _ in range(n))
        kb = 2 * k
        mask = (1 << kb) - 1
        km, keys = 0, []
        for i, c in enumerate(read):
            km = ((km << 2) | code3(c)) & mask
            if i + 1 >= k:
                keys.append(km)
        tops = []
        if n >= k:
            v = 0
            for i in range(n - k + 7):
                v = ((v << 2) | code3(read[i])) & 0x7FFF
                if i >= 6:
                    tops.append(v)
        assert len(tops) == len(keys)
        for d in range(1, 14):
            m = (1 << d) - 1
            assert [(x >> (kb - d)) & m for x in keys] == [(v >> (14 - d)) & m for v in tops]


def test_window_digit_needs_the_next_base_at_14_bits():
    """Bit 0 of the register (the low bit of the seventh slot) still lacks the
    next base's third code bit: at d = 14 the identity fails for a read whose
    eighth byte has that bit set, which is why the device stops at 13 bits."""
    k = 8
    read = b"AAAAAAAN"  # 'N' = 0x4E: code3 = 7, its third bit set
    km = 0
    for c in read:
        km = ((km << 2) | code3(c)) & ((1 << (2 * k)) - 1)
    v = 0
    for c in read[:7]:
        v = ((v << 2) | code3(c)) & 0x7FFF
    assert (km >> 2) & 0x3FFF != v & 0x3FFF
    assert (km >> 3) & 0x1FFF == (v >> 1) & 0x1FFF


def _segment_scatter(rng):
    d1, db = rng.randint(1, 3), rng.randint(1, 3)
    b1, ns = 1 << d1, 1 << db
    nblk, gf = rng.randint(1, 12), rng.randint(1, 2)
    spb = gf * rng.randint(1, 3)
    nsb = (nblk + spb - 1) // spb
    tq, grid = rng.randint(1, 6), rng.randint(1, 8)
    keys = []
    for b in range(nblk):
        for _ in range(rng.randint(0, 20)):
            keys.append((b, rng.randrange(b1), rng.randrange(ns), len(keys)))
    n = len(keys)
    h = np.zeros((b1, nblk), dtype=np.int64)
    h2 = np.zeros((b1 * ns, nsb), dtype=np.int64)
    for b, a, s, _ in keys:
        h[a, b] += 1
        h2[a * ns + s, b // spb] += 1
    o1 = np.concatenate([[0], np.cumsum(h.reshape(-1))])
    # level 1: scatter block sb = b // gf writes its keys of bucket a from
    # o1[a * nblk + sb * gf] on, in any order
    s1 = [None] * n
    cur = {}
    for b, a, s, i in keys:
        key = (a, b // gf)
        cur.setdefault(key, o1[a * nblk + (b // gf) * gf])
        s1[cur[key]] = (a, s, i)
        cur[key] += 1
    o2 = np.concatenate([[0], np.cumsum(h2.reshape(-1))])
    o2t = np.zeros(b1 * ns * nsb, dtype=np.int64)
    for a in range(b1):
        for g in range(nsb):
            for s in range(ns):
                o2t[(a * nsb + g) * ns + s] = o2[(a * ns + s) * nsb + g]

    def bounds(j):
        a, g = divmod(j, nsb)
        lo = g * spb
        return o1[a * nblk + lo], o1[a * nblk + min(lo + spb, nblk)]

    nseg = b1 * nsb

    def first_at(x):
        lo, hi = 0, nseg
        while lo < hi:
            mid = (lo + hi) // 2
            if bounds(mid)[0] < x:
                lo = mid + 1
            else:
                hi = mid
        return lo

    out = [None] * n
    for blk in range(grid):
        j = first_at(n * blk // grid)
        j1 = nseg if blk + 1 == grid else first_at(n * (blk + 1) // grid)
        for jj in range(j, j1):
            st, en = bounds(jj)
            base = [o2t[jj * ns + s] for s in range(ns)]
            for t0 in range(st, en, tq):
                tile = s1[t0:min(en, t0 + tq)]
                cnt = [0] * ns
                for a, s, i in tile:
                    p = base[s] + cnt[s]
                    cnt[s] += 1
                    assert out[p] is None
                    out[p] = (a, s, i)
                base = [base[s] + cnt[s] for s in range(ns)]
    assert all(x is not None for x in out)
    boff2 = [o2t[((g // ns) * nsb) * ns + g % ns] for g in range(b1 * ns)] + [n]
    for g in range(b1 * ns):
        for p in range(boff2[g], boff2[g + 1]):
            assert out[p][0] * ns + out[p][1] == g


def test_segment_scatter_places_every_key_in_its_bucket():
    rng = random.Random(11)
    for _ in range(200):
        _segment_scatter(rng)
