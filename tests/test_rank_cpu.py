"""Owner hashes and rank routing on the CPU: the restatements of
tests/rank_util.py pinned by known answers and closed forms, and the oracle
pinned by them (L-mer frequencies and min-L-mer routing on reads with
non-ACGT bytes, where the routing depends on bases outside the k-mer)."""
from __future__ import annotations

import numpy as np
import pytest

import rank_util as RU
from golden_util import read_fasta

NON_ACGT = (b">h\n" + b"ACGTNNacgtRYKMACGTTGCAAGGCTTAACCGGTTAA" * 6 + b"\n>h2\n"
            + b"nNnNACGTTTGGCCAAGGTTCCAAGGTTACGATCGATCGGATCAGCTAGC" * 4 + b"\n>h3\nACGNT\n")


def test_hash_known_answers():
    # worked by hand: a*x + b, then ((r >> 31) + r) & (2^31 - 1)
    assert RU.hash31(0) == 105929 and RU.uhash31(0) == 105929
    assert RU.hash31(1) == 68111 + 105929
    # 68111 * 2^63 wraps to -2^63 (signed): the high terms have no low 31 bits
    assert RU.hash31(1 << 63) == 105929
    # x = -1: a*x + b = 105929 - 68111 = 37818 in either width
    assert RU.hash31((1 << 64) - 1) == 37818 and RU.uhash31((1 << 128) - 1) == 37818
    # 128-bit: 68111 * 2^64 + 105929, shifted terms are multiples of 2^33
    assert RU.uhash31(1 << 64) == 105929
    # a value whose 64-bit product goes negative: ((v >> 31) + v) uses the
    # arithmetic shift; compare with two's-complement arithmetic done by numpy
    x = 0x9E3779B97F4A7C15
    with np.errstate(over="ignore"):
        v = (np.uint64(68111) * np.uint64(x) + np.uint64(105929)).astype(np.int64)
        h = ((v >> np.int64(31)) + v) & np.int64((1 << 31) - 1)
    assert RU.hash31(x) == int(h)


def test_hash31_matches_oracle_lmer_owner(oracle_mod):
    rng = np.random.default_rng(5)
    for x in rng.integers(0, 1 << 16, 200):
        for P in (2, 3, 7, 8):
            assert RU.hash31(int(x)) % P == oracle_mod.lmer_owner(int(x), P)


@pytest.mark.parametrize("fa", [NON_ACGT, None])
def test_lmer_frequency_closed_form(oracle_mod, fa):
    fa = fa if fa is not None else read_fasta("syn_tiny.fa.gz")
    assert np.array_equal(RU.lmer_freq(fa), oracle_mod.lmer_frequency(fa))


def test_owner_walk_is_leftmost_minimum_on_acgt():
    """On ACGT-only reads the restated routing is the per-window closed form:
    hash31 of the leftmost minimum-frequency 8-mer of the window."""
    fa = read_fasta("syn_tiny.fa.gz")[:20000]
    fa = fa[:fa.rfind(b"\n>") + 1]
    freq = RU.lmer_freq(fa)
    for k in (8, 21, 32):
        walk = RU.owner_walk(fa, k, 5, freq)
        got = iter(walk)
        for line in RU._records(fa, k):
            c = [RU.code3(b) for b in line]
            lm = []
            v = 0
            for x in c:
                v = ((v << 2) | x) & RU.LMASK
                lm.append(v)
            for i in range(k - 1, len(c)):
                if k == RU.L and i == k - 1:
                    own = RU.hash31(0) % 5  # the first window keeps the initial L-mer 0
                else:
                    ls = lm[i - k + RU.L:i + 1]
                    fs = [int(freq[x]) for x in ls]
                    own = RU.hash31(ls[fs.index(min(fs))]) % 5
                assert next(got)[1] == own


@pytest.mark.parametrize("k,P", [(8, 3), (21, 3), (32, 2), (33, 5), (64, 3)])
def test_oracle_rank_share_matches_owner_walk(oracle_mod, k, P):
    freq = RU.lmer_freq(NON_ACGT)
    walk = RU.owner_walk(NON_ACGT, k, P, freq)
    for r in range(P):
        keys, counts = oracle_mod.count(NON_ACGT, k, nranks=P, rank=r)
        assert RU.table_counter(keys, counts) == RU.share_counter(walk, r), r


@pytest.mark.parametrize("name,k,t", [("syn_mid.fa.gz", 32, 21), ("syn_rep.fa.gz", 64, 15), ("syn_tiny.fa.gz", 33, 21)])
def test_pipeline_at_p_ranks(oracle_mod, name, k, t):
    """The CPU baseline's multi-rank pipeline (the reference at P ranks, one
    thread per rank) gives the one-rank totals on ACGT-only reads: the ranks'
    tables partition the table, the histogram sums to the one-rank
    histogram, and every node is built by exactly its owner."""
    fa = read_fasta(name)
    keys, counts = oracle_mod.count(fa, k)
    hist, mb, keep, rc = oracle_mod.cutoff(counts, t)
    nodes = oracle_mod.macro_nodes(keys[keep], counts[keep], k)
    for P in (1, 2, 5):
        r = oracle_mod.pipeline_ranks(fa, k, t, P)
        assert (r["distinct"], r["kept"], r["nodes"], r["min_bucket"]) == (
            keys.shape[0], int(keep.sum()), nodes["keys"].shape[0], mb), P
        assert np.array_equal(r["hist"], hist) and r["instances"] == int(counts.sum())
