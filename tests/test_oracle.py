"""The CPU oracle pinned against the reference's golden vectors, the reference's
own cutoff / macro-node code (tests/golden/*.ref.txt*, made by
oracle/_ref from the reference sources), a brute-force counter and closed forms."""
from __future__ import annotations

import numpy as np
import pytest

from golden_util import SYNTH, read_fasta, read_ref, ref_nodes_as_arrays

pytestmark = pytest.mark.filterwarnings("ignore")


def _words(k):
    return 1 if k <= 32 else 2


def _kmer_sum_expected(fa: bytes, k: int) -> int:
    """sum over read lines of max(0, len-k+1) (strict single-line FASTA)."""
    lines = fa.split(b"\n")
    tot = 0
    for i in range(0, len(lines) - 1, 2):
        assert lines[i].startswith(b">")
        tot += max(0, len(lines[i + 1]) - k + 1)
    return tot


def test_reference_unit_test_invariant(oracle_mod):
    # src/concurrent_kmer_counter_test.cpp:96,113: sum(freq) == reads*(100-k+1) == 138
    fa = read_fasta("ref_test_two_reads.fa")
    keys, counts = oracle_mod.count(fa, 32)
    assert int(counts.sum()) == 138
    reads, bkeys, bcounts = oracle_mod.builder_count(fa, 32, 100, 64)
    assert reads == 2 and int(bcounts.sum()) == 138
    # on ACGT-only 100 bp reads the builder API and the main path agree
    assert np.array_equal(bkeys, keys) and np.array_equal(bcounts, counts)


@pytest.mark.parametrize("name", ["syn_tiny.fa.gz", "syn_rep.fa.gz", "read_test1.fa"])
@pytest.mark.parametrize("k", [8, 21, 31, 32, 33, 47, 63, 64])
def test_count_matches_bruteforce(oracle_mod, name, k):
    fa = read_fasta(name)
    keys, counts = oracle_mod.count(fa, k)
    bk, bc = oracle_mod.bruteforce_count(fa, k)
    assert keys.shape[1] == _words(k)
    assert np.array_equal(keys, bk) and np.array_equal(counts, bc)
    assert int(counts.sum()) == _kmer_sum_expected(fa, k)
    # strictly ascending unsigned keys, counts >= 1
    if _words(k) == 1:
        assert np.all(keys[1:, 0] > keys[:-1, 0])
    else:
        hi, lo = keys[:, 1], keys[:, 0]
        assert np.all((hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1]) & (lo[1:] > lo[:-1])))
    assert counts.min() >= 1


def test_closed_form_single_window(oracle_mod):
    # one read of exactly k bases -> one k-mer whose value is the 2-bit packing
    # with A=0 C=1 T=2 G=3 (distribute_kmers.h:784-786), first base most significant
    seq = "ACGTTGCAAGCTTAGCGATCGATCGGCTAGCT"
    fa = (">x\n" + seq + "\n").encode()
    keys, counts = oracle_mod.count(fa, 32)
    code = {"A": 0, "C": 1, "T": 2, "G": 3}
    v = 0
    for c in seq:
        v = (v << 2) | code[c]
    assert keys.shape == (1, 1) and int(keys[0, 0]) == v and counts[0] == 1


def test_non_acgt_three_bit_or(oracle_mod):
    # 'N' has 3-bit code 7: its bit 2 ORs into the previous base's low bit
    # (kmer_shift, distribute_kmers.h:798-805).  Lower case maps like upper case.
    k = 8
    fa = b">a\nACGTNacgtn\n"
    keys, counts = oracle_mod.count(fa, k)
    bk, bc = oracle_mod.bruteforce_count(fa, k)
    assert np.array_equal(keys, bk) and np.array_equal(counts, bc)
    # closed form of the first window "ACGTNacg"
    codes = [((ord(c) >> 1) & 7) for c in "ACGTNacg"]
    v = 0
    for j, c in enumerate(codes):
        v |= c << (2 * (k - 1 - j))
    v &= (1 << (2 * k)) - 1
    assert v in set(int(x) for x in keys[:, 0])
    # a plain 2-bit packing would differ: the T before N becomes T|1 = G
    plain = 0
    for c in "ACGTNacg":
        plain = (plain << 2) | (((ord(c) >> 1) & 7) & 3)
    assert plain != v


@pytest.mark.parametrize("k", [32, 64])
def test_batches_and_aggregate_mode_do_not_change_table(oracle_mod, k):
    fa = read_fasta("syn_tiny.fa.gz")
    keys, counts = oracle_mod.count(fa, k)
    for batch, inplace in ((20000, False), (5000, True), (300000, True)):
        kb, cb = oracle_mod.count(fa, k, batch=batch, inplace=inplace)
        assert np.array_equal(kb, keys) and np.array_equal(cb, counts)


@pytest.mark.parametrize("k", [32, 33])
@pytest.mark.parametrize("P", [2, 3, 7])
def test_rank_shares_partition_the_table(oracle_mod, k, P):
    fa = read_fasta("syn_tiny.fa.gz")
    keys, counts = oracle_mod.count(fa, k)
    allk, allc = [], []
    for r in range(P):
        kr, cr = oracle_mod.count(fa, k, nranks=P, rank=r)
        allk.append(kr)
        allc.append(cr)
    K = np.concatenate(allk)
    C = np.concatenate(allc)
    assert K.shape[0] == keys.shape[0]  # shares are disjoint and cover the table
    order = np.lexsort(K.T) if K.shape[1] == 2 else np.argsort(K[:, 0], kind="stable")
    assert np.array_equal(K[order], keys) and np.array_equal(C[order], counts)


def test_hash31_closed_form(oracle_mod):
    # hash31 (distribute_kmers.h:856-873) on small values has no wrap-around:
    # ((a*x+b) + ((a*x+b) >> 31)) & (2^31-1), then % P
    a, b, M = 68111, 105929, (1 << 31) - 1
    for x in (0, 1, 2, 77, 4095, 65535):
        r = a * x + b
        h = ((r >> 31) + r) & M
        for P in (1, 2, 3, 8, 64):
            assert oracle_mod.lmer_owner(x, P) == h % P


def test_lmer_frequency_total(oracle_mod):
    fa = read_fasta("syn_tiny.fa.gz")
    f = oracle_mod.lmer_frequency(fa)
    assert int(f.sum()) == _kmer_sum_expected(fa, 8)


def test_format_errors(oracle_mod):
    # READ_TEST1 as written in the reference test ends with a blank line, which
    # the main path's parser rejects (assert at src/kmer_counting.cpp:2928)
    raw = read_fasta("read_test1_raw.fa")
    with pytest.raises(oracle_mod.OracleError) as e:
        oracle_mod.count(raw, 32)
    assert e.value.code == oracle_mod.PK_EFORMAT
    # a wrapped sequence (two sequence lines) is rejected too
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.count(b">a\n" + b"A" * 40 + b"\n" + b"C" * 40 + b"\n>b\n" + b"G" * 40 + b"\n", 32)
    # a record whose header ends fewer than k bytes before the end stops the
    # scan (break at :2936), so nothing after it is checked
    keys, counts = oracle_mod.count(b">a\n" + b"ACGT" * 10 + b"\n>b\nx\nyy", 32)
    assert int(counts.sum()) == 9
    # bytes other than '>' right after a sequence line are rejected
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.count(b">a\n" + b"ACGT" * 10 + b"\nxx", 32)


def test_cutoff_closed_forms(oracle_mod):
    # first argmin over hist[1..t-1] (kmer_histogram.cpp:114-115)
    counts = np.array([1, 1, 1, 2, 3, 3, 5, 5, 5, 5, 40], dtype=np.int32)
    hist, mb, keep, rc = oracle_mod.cutoff(counts, 6)
    assert list(hist) == [0, 3, 1, 2, 0, 4] and mb == 4 and rc == 0
    assert np.array_equal(keep, counts >= 4)
    # ties -> the first index
    hist, mb, keep, rc = oracle_mod.cutoff(np.array([1, 2, 3, 9], dtype=np.int32), 5)
    assert mb == 4
    # t == 1 -> empty range -> min_bucket 1 (keep everything)
    hist, mb, keep, rc = oracle_mod.cutoff(np.array([1, 2], dtype=np.int32), 1)
    assert mb == 1 and keep.all()
    # every k-mer seen once -> min_bucket 2 -> nothing kept -> the reference aborts
    hist, mb, keep, rc = oracle_mod.cutoff(np.ones(10, dtype=np.int32), 21)
    assert mb == 2 and rc == oracle_mod.PK_EEMPTY


def _golden_cases():
    for name, (cov, ks, t) in SYNTH.items():
        for k in ks:
            yield name, k, cov, t


@pytest.mark.parametrize("name,k,cov,t", list(_golden_cases()))
def test_oracle_matches_reference_pipeline(oracle_mod, name, k, cov, t):
    """Counting stats, cutoff, macro nodes (prefix order included) and visit
    values against the reference's own kmer_histogram.cpp / mn_construct.cpp /
    wire.cpp run on the same input."""
    fa = read_fasta(f"syn_{name}.fa.gz")
    ref = read_ref(f"syn_{name}_k{k}.ref.txt.gz")
    keys, counts = oracle_mod.count(fa, k)
    assert keys.shape[0] == ref.distinct
    hist, mb, keep, rc = oracle_mod.cutoff(counts, t)
    assert rc == 0 and mb == ref.min_bucket and int(keep.sum()) == ref.kept
    nodes = oracle_mod.macro_nodes(keys[keep], counts[keep], k, ref_order=True)
    want = ref_nodes_as_arrays(ref, _words(k))
    assert nodes["keys"].shape[0] == ref.macro_nodes == want["keys"].shape[0]
    for f in ("keys", "pre_cnt", "suf_cnt", "pre_order", "suf_order"):
        assert np.array_equal(nodes[f], want[f]), f
    for c, v in want["visits"][:2000]:
        assert oracle_mod.visit_value(c, cov) == v
    # canonical order: same lists, prefixes ascending by base code
    canon = oracle_mod.macro_nodes(keys[keep], counts[keep], k, ref_order=False)
    for f in ("keys", "pre_cnt", "suf_cnt", "suf_order"):
        assert np.array_equal(canon[f], want[f])
    po = canon["pre_order"].astype(np.int32)
    po[po == 255] = 99
    assert np.all((np.diff(po, axis=1) > 0) | (po[:, 1:] == 99))


def test_reference_pipeline_detects_planted_mistakes(oracle_mod):
    """The golden comparison fails for plausible mistakes planted in the oracle's
    output: a transposed prefix order, an off-by-one count, a dropped node."""
    fa = read_fasta("syn_rep.fa.gz")
    ref = read_ref("syn_rep_k32.ref.txt.gz")
    keys, counts = oracle_mod.count(fa, 32)
    hist, mb, keep, rc = oracle_mod.cutoff(counts, 15)
    nodes = oracle_mod.macro_nodes(keys[keep], counts[keep], 32, ref_order=True)
    want = ref_nodes_as_arrays(ref, 1)
    multi = np.where((nodes["pre_order"][:, 1] != 255))[0]
    assert multi.size > 0
    po = nodes["pre_order"].copy()
    po[multi[0], [0, 1]] = po[multi[0], [1, 0]]
    assert not np.array_equal(po, want["pre_order"])
    sc = nodes["suf_cnt"].copy()
    sc[np.nonzero(sc)[0][0], np.nonzero(sc)[1][0]] += 1
    assert not np.array_equal(sc, want["suf_cnt"])
    assert nodes["keys"][1:].shape != want["keys"].shape
    # the reference's cutoff abort case (all k-mers unique)
    two = read_ref("ref_test_two_reads_k32.ref.txt")
    assert two.aborted and two.min_bucket == 2 and two.kept == 0 and two.distinct == 138


# ---------------------------------------------------------------------------
# Builder API (concurrent_kmer_counter.hpp:2018-2082): the oracle against an
# independent closed-form restatement (tests/builder_util.py).

from builder_util import builder_closed_form, edge_inputs, table_to_counter  # noqa: E402

BUILDER_PARAMS = [(1, 100, 64), (5, 30, 64), (21, 100, 64), (32, 100, 64), (33, 100, 64),
                  (32, 100, 128), (40, 60, 128), (64, 100, 128)]


@pytest.mark.parametrize("name", sorted(edge_inputs()))
@pytest.mark.parametrize("klen,read_len,bits", BUILDER_PARAMS)
def test_builder_oracle_matches_closed_form(oracle_mod, name, klen, read_len, bits):
    fa = edge_inputs()[name]
    reads, keys, counts = oracle_mod.builder_count(fa, klen, read_len, bits)
    want_reads, want = builder_closed_form(fa, klen, read_len, bits)
    assert reads == want_reads
    assert table_to_counter(keys, counts) == want
    if keys.shape[0]:
        assert keys.shape[1] == bits // 64


def test_builder_oracle_custom_symbols(oracle_mod):
    fa = edge_inputs()["ragged"].replace(b"\n", b"|").replace(b">", b"@")
    reads, keys, counts = oracle_mod.builder_count(fa, 21, 100, 64, b"|", b"@")
    want_reads, want = builder_closed_form(fa, 21, 100, 64, b"|", b"@")
    assert reads == want_reads and reads > 0
    assert table_to_counter(keys, counts) == want


def test_builder_closed_form_detects_planted_mistakes():
    """The restatement is sensitive to the rules it checks: a wrong base code,
    a dropped short-read rule or a wrong record rule change its result."""
    fa = edge_inputs()["junk_and_multi_gt"]
    reads, want = builder_closed_form(fa, 21, 100, 64)
    _, alt = builder_closed_form(fa.replace(b"G", b"T"), 21, 100, 64)
    assert alt != want
    # reads shorter than klen are skipped: read_len 20 < klen 21 leaves none
    assert builder_closed_form(fa, 21, 20, 64)[0] == 0 and builder_closed_form(fa, 20, 20, 64)[0] > 0
    # the first '>' of a sequence line starts a record
    r3, _ = builder_closed_form(fa.replace(b"seq>with>gt", b"seq with gt"), 21, 100, 64)
    assert r3 == reads - 1


def test_rank_share_owner_closed_form(oracle_mod):
    """On ACGT-only reads the owner the oracle tracks incrementally equals the
    closed form: hash31 of the leftmost minimum-frequency 8-mer of the k-mer
    (frequencies = numpy count of every 8-mer of every read), modulo P."""
    from collections import Counter
    from numpy.lib.stride_tricks import sliding_window_view
    fa = read_fasta("syn_tiny.fa.gz")
    k, L, P = 32, 8, 3
    code = np.zeros(256, dtype=np.uint64)
    for ch, v in zip(b"ACTG", range(4)):
        code[ch] = v
    reads = [np.frombuffer(ln, dtype=np.uint8) for ln in fa.split(b"\n")[1::2] if len(ln) >= L]
    lm_all = []
    per_read = []
    for r in reads:
        c = code[r]
        lm = np.zeros(len(c) - L + 1, dtype=np.uint64)
        for j in range(L):
            lm = (lm << np.uint64(2)) | c[j:len(c) - L + 1 + j]
        lm_all.append(lm)
        per_read.append((c, lm))
    freq = np.bincount(np.concatenate(lm_all).astype(np.int64), minlength=1 << 16)
    assert np.array_equal(freq.astype(np.uint64), oracle_mod.lmer_frequency(fa))
    a, b, M = 68111, 105929, (1 << 31) - 1
    hv = a * np.arange(1 << 16, dtype=np.int64) + b
    owner_of = ((hv >> 31) + hv) & M
    want = [Counter() for _ in range(P)]
    for c, lm in per_read:
        if len(c) < k:
            continue
        win = sliding_window_view(freq[lm.astype(np.int64)], k - L + 1)  # per k-mer
        first = np.argmin(win, axis=1)                                  # leftmost minimum
        mins = lm[np.arange(win.shape[0]) + first]
        own = owner_of[mins.astype(np.int64)] % P
        km = np.zeros(len(c) - k + 1, dtype=np.uint64)
        for j in range(k):
            km = (km << np.uint64(2)) | c[j:len(c) - k + 1 + j]
        for x, o in zip(km.tolist(), own.tolist()):
            want[o][x] += 1
    for r in range(P):
        kr, cr = oracle_mod.count(fa, k, nranks=P, rank=r)
        got = Counter({int(x): int(y) for x, y in zip(kr[:, 0], cr)})
        assert got == want[r], r
