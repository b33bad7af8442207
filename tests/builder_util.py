"""Inputs and an independent restatement for the builder-API counter
(Concurrent_Kmer_Partial_Counter::default_count_and_process,
src/concurrent_kmer_counter.hpp:2018-2082).

builder_closed_form() finds the records by their characterisation (the
first skip symbol after every delimiter, or of the buffer) rather than by
the reference's forward scan, and computes every key as the base-4 number of
the last W = bits/2 symbols shifted in since the read start (closed form),
rather than by the rolling update.  windows_closed_form() does the same for
the implicit counters' raw-byte windows (src/concurrent_kmer_counter.hpp:
910-1022,1363-1578).  Pure Python: small inputs only."""
from __future__ import annotations

from collections import Counter

import numpy as np

SYM = {ord("A"): 0, ord("C"): 1, ord("T"): 2, ord("G"): 3}


def builder_closed_form(fa: bytes, klen: int, read_len: int, bits: int,
                        delimiter: bytes = b"\n", skip: bytes = b">"):
    d, s = delimiter[0], skip[0]
    ev = [i for i, b in enumerate(fa) if b == d or b == s]
    counts: Counter = Counter()
    reads = 0
    W = bits // 2
    for j, p in enumerate(ev):
        if fa[p] != s or (j > 0 and fa[ev[j - 1]] == s):
            continue
        hdr = next((q for q in ev[j + 1:] if fa[q] == d), None)
        if hdr is None:
            break
        rb = hdr + 1
        rl = min(read_len, len(fa) - rb)
        if rl < klen:
            continue
        reads += 1
        codes = [SYM.get(b, 0) for b in fa[rb:rb + rl]]
        for i in range(rl - klen + 1):
            end = i + klen
            key = 0
            for c in codes[max(0, end - W):end]:
                key = key * 4 + c
            counts[key] += 1
    return reads, counts


def windows_closed_form(fa: bytes, klen: int, read_len: int, mode: str,
                        delimiter: bytes = b"\n", skip: bytes = b">"):
    """(records, windows, distinct) of the implicit builder counters
    (omp_implicit_v0 / _v1): the records as above but every one counted,
    each read's klen-byte windows as raw byte strings, two windows equal when
    strncmp(a, b, klen) == 0, i.e. their bytes agree up to and including the
    first NUL.  Raises ValueError where the device reports PK_EFORMAT: v0 on
    a read shorter than read_len, v1 on a read shorter than klen."""
    d, s = delimiter[0], skip[0]
    ev = [i for i, b in enumerate(fa) if b == d or b == s]
    reads, windows, seen = 0, 0, set()
    for j, p in enumerate(ev):
        if fa[p] != s or (j > 0 and fa[ev[j - 1]] == s):
            continue
        hdr = next((q for q in ev[j + 1:] if fa[q] == d), None)
        if hdr is None:
            break
        rb = hdr + 1
        rl = min(read_len, len(fa) - rb)
        if mode == "v0" and rl != read_len:
            raise ValueError("Variable size read")
        if mode == "v1" and rl < klen:
            raise ValueError("read shorter than klen")
        reads += 1
        for i in range(rl - klen + 1):
            w = fa[rb + i:rb + i + klen]
            z = w.find(b"\0")
            seen.add(w if z < 0 else w[:z + 1])
            windows += 1
    return reads, windows, len(seen)


def table_to_counter(keys: np.ndarray, counts: np.ndarray) -> Counter:
    out: Counter = Counter()
    for i in range(keys.shape[0]):
        v = int(keys[i, 0])
        if keys.shape[1] == 2:
            v |= int(keys[i, 1]) << 64
        out[v] = int(counts[i])
    return out


def edge_inputs() -> dict:
    """Small buffers exercising the builder's record rules."""
    rng = np.random.default_rng(11)
    acgt = np.frombuffer(b"ACGT", dtype=np.uint8)
    parts = []
    for i in range(60):
        L = int(rng.integers(20, 160))
        seq = bytearray(acgt[rng.integers(0, 4, L)].tobytes())
        for j in rng.integers(0, L, 3):
            seq[j] = b"NacgR"[int(rng.integers(0, 5))]
        parts.append(b">read%d some text with ACGT letters\n" % i + bytes(seq) + b"\n")
    ragged = b"".join(parts)
    return {
        "ragged": ragged,
        "junk_and_multi_gt": (b"junk line\nmore>junk\n>h1>also>header\nACGTACGTNNacgtACGTACGTTTGGCCAAGG\n"
                              b">h2\nGGGGCCCCAAAATTTTACGTACGTACGTAC\n>short\nACG\n"
                              b"seq>with>gt\nTTTTGGGGCCCCAAAATTTTGGGGCCCCAAAA\n>r3\n"
                              b"ACGTTGCAACGTTGCAACGTTGCAACGTTGCAAC"),
        "header_without_newline": b">a\n" + b"ACGT" * 30 + b"\n>b" + b"G" * 40,
        "empty_reads": b">a\n\n>b\n\n>c\n" + b"C" * 50 + b"\n",
        "no_records": b"ACGTACGTACGT\nACGT\n",
        "empty": b"",
    }
