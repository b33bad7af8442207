"""Restatements for the rank-share checks, written from the reference's
formulas (paths of the vkundeti/pakman_plus tree):

  hash31 / uhash31     includes/distribute_kmers.h:836-873, retrieve_proc_id
                       src/distribute_kmers.cpp:95-120: (a*x + b), then
                       ((r >> 31) + r) & (2^31 - 1); hash31 in signed 64-bit
                       arithmetic (kmer_t = uint64_t builds, k <= 32), uhash31
                       in unsigned 128-bit (EXTEND_KMER builds, k > 32); owner =
                       that value % P.  Python integers with explicit wraps.
  lmer_freq            Sliding_window_l, src/kmer_counting.cpp:391-438, as a
                       closed form per position: the L-mer ending at byte e of a
                       read line is (OR_{j<8} code3(c[e-7+j]) << 2(7-j)) & 0xFFFF
                       (the rolling 3-bit update: a byte's third bit lands in
                       the next-older slot and falls out of the 16-bit mask
                       after eight shifts).
  owner_walk           Sliding_window's owner routing, src/kmer_counting.cpp:
                       2920-3050 with recalculate_min_lmer :355-388, restated
                       window by window: the minimum carried from the previous
                       window while it stays inside, else the leftmost minimum
                       of the L-mers read back from the packed k-mer.  With
                       non-ACGT bytes the read-back L-mers differ from the
                       rolling ones (the next byte's third bit ORs into their
                       last slot), so the routing depends on that history and
                       is restated, not derived per window; on ACGT-only reads
                       it is the leftmost minimum of each window's L-mers
                       (checked against that closed form).

Pure Python / numpy, small inputs only."""
from __future__ import annotations

from collections import Counter

import numpy as np

A, B, MOD = 68111, 105929, (1 << 31) - 1
M64, M128 = (1 << 64) - 1, (1 << 128) - 1
L = 8
LMASK = (1 << (2 * L)) - 1


def _signed64(v: int) -> int:
    v &= M64
    return v - (1 << 64) if v >> 63 else v


def hash31(x: int) -> int:
    v = _signed64(A * _signed64(x) + B)
    return _signed64((v >> 31) + v) & MOD


def uhash31(x: int) -> int:
    r = (A * x + B) & M128
    r = ((r >> 31) + r) & M128
    return r & MOD


def key_owner(x: int, k: int, P: int) -> int:
    """retrieve_proc_id(kmer_t) % P for a k-mer (or node key) of a k-build."""
    return (hash31(x) if k <= 32 else uhash31(x)) % P


def code3(b: int) -> int:
    return (b >> 1) & 7


def _records(fa: bytes, need: int):
    """Read lines of the main path's record scan (kmer_counting.cpp:2920-2945;
    Sliding_window_l uses need = L, Sliding_window need = k)."""
    p = fa.find(b">")
    if p < 0:
        return
    n = len(fa)
    while p < n:
        assert fa[p] == ord(">")
        q = fa.find(b"\n", p)
        p = n if q < 0 else q + 1
        if p + need > n:
            return
        q = fa.find(b"\n", p)
        e = n if q < 0 else q
        yield fa[p:e]
        p = e + 1


def lmer_freq(fa: bytes) -> np.ndarray:
    freq = np.zeros(1 << (2 * L), dtype=np.uint64)
    for line in _records(fa, L):
        c = (np.frombuffer(line, dtype=np.uint8) >> 1) & 7
        if c.size < L:
            continue
        v = np.zeros(c.size - L + 1, dtype=np.int64)
        for j in range(L):
            v |= c[j:c.size - L + 1 + j].astype(np.int64) << (2 * (L - 1 - j))
        np.add.at(freq, v & LMASK, 1)
    return freq


def owner_walk(fa: bytes, k: int, P: int, freq: np.ndarray) -> list:
    """[(k-mer, owner)] in input order for every window of every read.  State
    per read: the current minimum (frequency, L-mer) and its start relative to
    the next window, all zero at the read start (so with k = 8 the first
    window keeps L-mer 0, as the reference does)."""
    kmask = (1 << (2 * k)) - 1
    out = []
    for line in _records(fa, k):
        codes = [code3(b) for b in line]
        if len(codes) < k:
            continue
        km, lm = 0, 0
        min_f, min_l, min_pos = 0, 0, 0
        for i in range(k - 1):
            km = ((km << 2) | codes[i]) & kmask
            lm = ((lm << 2) | codes[i]) & LMASK
            if i >= L - 1:
                f = int(freq[lm])
                if i == L - 1:
                    min_f, min_l = f, lm
                elif f < min_f:
                    min_f, min_l, min_pos = f, lm, i - (L - 1)
        for i in range(k - 1, len(codes)):
            km = ((km << 2) | codes[i]) & kmask
            lm = ((lm << 2) | codes[i]) & LMASK
            if min_pos < 0:
                # the minimum left the window: leftmost minimum of the L-mers
                # read back from the packed k-mer (2-bit fields)
                for kk in range(k - L + 1):
                    v = (km >> (2 * (k - kk - L))) & LMASK
                    f = int(freq[v])
                    if kk == 0 or f < min_f:
                        min_f, min_l, min_pos = f, v, kk
            f = int(freq[lm])
            if f < min_f:
                min_f, min_l, min_pos = f, lm, k - L
            min_pos -= 1
            out.append((km, hash31(min_l) % P))
    return out


def share_counter(walk: list, rank: int) -> Counter:
    c: Counter = Counter()
    for km, o in walk:
        if o == rank:
            c[km] += 1
    return c


def table_counter(keys: np.ndarray, counts: np.ndarray) -> Counter:
    out: Counter = Counter()
    for i in range(keys.shape[0]):
        v = int(keys[i, 0]) | ((int(keys[i, 1]) << 64) if keys.shape[1] == 2 else 0)
        out[v] = int(counts[i])
    return out
