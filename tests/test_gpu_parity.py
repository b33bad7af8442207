"""HIP library (through its C ABI) against the CPU oracle and the reference's
own cutoff / macro-node output.  Bit-exact: keys, counts, histogram,
min_bucket, kept table and macro-node records."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from golden_util import SYNTH, read_fasta, read_ref, ref_nodes_as_arrays

pytestmark = pytest.mark.gpu


def _p(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from pakman_plus_amd import api
    c = api.Context(0)
    yield c
    c.close()


def _words(k):
    return 1 if k <= 32 else 2


def _check_table(ctx, oracle_mod, fa, k):
    ctx.count(fa, k)
    keys, counts = ctx.table_host()
    ok, oc = oracle_mod.count(fa, k)
    assert keys.shape == ok.shape, (keys.shape, ok.shape)
    assert np.array_equal(keys, ok)
    assert np.array_equal(counts, oc)
    return keys, counts


def _check_nodes(ctx, oracle_mod, k, kept_keys, kept_counts):
    from pakman_plus_amd import api
    m = ctx.macro_nodes()
    got = api.nodes_as_arrays(ctx.nodes_host(), k)
    want = oracle_mod.macro_nodes(kept_keys, kept_counts, k)
    assert m == want["keys"].shape[0]
    for f in ("keys", "pre_cnt", "suf_cnt"):
        assert np.array_equal(got[f], want[f]), f
    return got


def test_reference_unit_test_input(ctx, oracle_mod):
    _p("case: ref_test_two_reads k=32")
    from pakman_plus_amd import api
    fa = read_fasta("ref_test_two_reads.fa")
    keys, counts = _check_table(ctx, oracle_mod, fa, 32)
    assert int(counts.sum()) == 138  # src/concurrent_kmer_counter_test.cpp:96
    with pytest.raises(api.PakHipError) as e:
        ctx.cutoff(21)
    assert e.value.code == api.PK_EEMPTY and ctx.min_bucket == 2


@pytest.mark.parametrize("name,k,cov,t", [(n, k, c, t) for n, (c, ks, t) in SYNTH.items() for k in ks])
def test_golden_pipeline(ctx, oracle_mod, name, k, cov, t):
    _p(f"case: golden {name} k={k}")
    fa = read_fasta(f"syn_{name}.fa.gz")
    ref = read_ref(f"syn_{name}_k{k}.ref.txt.gz")
    keys, counts = _check_table(ctx, oracle_mod, fa, k)
    assert keys.shape[0] == ref.distinct
    hist, mb, nk = ctx.cutoff(t)
    ohist, omb, okeep, _ = oracle_mod.cutoff(counts, t)
    assert np.array_equal(hist, ohist) and mb == omb == ref.min_bucket and nk == ref.kept
    kk, kc = ctx.kept_host()
    assert np.array_equal(kk, keys[okeep]) and np.array_equal(kc, counts[okeep])
    got = _check_nodes(ctx, oracle_mod, k, kk, kc)
    want = ref_nodes_as_arrays(ref, _words(k))
    for f in ("keys", "pre_cnt", "suf_cnt"):
        assert np.array_equal(got[f], want[f]), f


@pytest.mark.parametrize("k", [8, 9, 15, 16, 17, 24, 31, 32, 33, 40, 48, 63, 64])
def test_count_all_k(ctx, oracle_mod, k):
    _p(f"case: tiny all-k k={k}")
    fa = read_fasta("syn_tiny.fa.gz")
    _check_table(ctx, oracle_mod, fa, k)


def test_format_errors(ctx, oracle_mod):
    _p("case: format errors")
    from pakman_plus_amd import api
    with pytest.raises(api.PakHipError) as e:
        ctx.count(read_fasta("read_test1_raw.fa"), 32)
    assert e.value.code == api.PK_EFORMAT
    with pytest.raises(api.PakHipError) as e:
        ctx.count(b">a\n" + b"A" * 40 + b"\n" + b"C" * 40 + b"\n>b\n" + b"G" * 40 + b"\n", 32)
    assert e.value.code == api.PK_EFORMAT
    with pytest.raises(api.PakHipError) as e:
        ctx.count(b">a\n" + b"ACGT" * 10 + b"\nxx", 32)
    assert e.value.code == api.PK_EFORMAT
    with pytest.raises(api.PakHipError) as e:
        ctx.count(b">a\nACGT\n", 7)
    assert e.value.code == api.PK_EINVAL
    # a tail too short to hold a k-mer is never inspected
    _check_table(ctx, oracle_mod, b">a\n" + b"ACGT" * 10 + b"\n>b\nx\nyy", 32)
    # the fixed READ_TEST1 parses
    _check_table(ctx, oracle_mod, read_fasta("read_test1.fa"), 32)


@pytest.mark.parametrize("fa", [
    b"",
    b"no header at all\n",
    b">only header",
    b">h\nACGT\n",
    b"junk before\n>h\n" + b"ACGTTGCA" * 8 + b"\n",
    b">h\n" + b"ACGTNNacgtRYKM" * 9 + b"\n>h2\n" + b"nNnNACGT" * 7,
    b">h\r\n" + b"ACGT" * 20 + b"\n",
    # no '>' in the line index's first 16 KB block: the first one is found
    # by a second pass over every block (and none at all gives no records)
    b"junk\n" * 5000 + b">h\n" + b"ACGTTGCA" * 8 + b"\n>h2\n" + b"TTGCAACG" * 9 + b"\n",
    b"junk line without a record\n" * 2000,
])
def test_edge_inputs(ctx, oracle_mod, fa):
    _p(f"case: edge input len={len(fa)}")
    for k in (8, 32, 33, 64):
        _check_table(ctx, oracle_mod, fa, k)


def test_mixed_read_lengths(ctx, oracle_mod):
    """Reads from 20 bp to 150 kbp in one input: the level-1 chunks stage
    their bytes in LDS sized for the mean record, and a read longer than the
    staging buffer is walked from global memory, alone in its chunk."""
    _p("case: mixed read lengths")
    rng = np.random.default_rng(11)
    lens = [20, 150, 40_000, 90, 150_000, 33, 64, 65, 2_000, 100_000] + [100] * 300 + [70_000]
    recs = []
    for i, n in enumerate(lens):
        seq = rng.choice(np.frombuffer(b"ACGTACGTN", dtype=np.uint8), size=n)
        recs.append(b">r%d\n" % i + seq.tobytes() + b"\n")
    fa = b"".join(recs)
    for k in (21, 32, 64):
        _check_table(ctx, oracle_mod, fa, k)


def test_host_input_paths(ctx, gen_mod):
    """A 33 MB host FASTA goes through the pinned staging buffers when it is
    pageable and is copied directly when it is pinned or already on the
    device: the three tables are identical."""
    import torch
    _p("case: host input paths")
    fa = gen_mod.make_fasta(gen_mod.config_small())
    assert fa.nbytes > (16 << 20)
    ctx.count(fa, 32)
    k1, c1 = ctx.table_host()
    pinned = torch.empty(fa.nbytes, dtype=torch.uint8, pin_memory=True)
    pinned.numpy()[:] = fa
    ctx.count(pinned.numpy(), 32)
    k2, c2 = ctx.table_host()
    dev = torch.from_numpy(fa).to("cuda:0")
    torch.cuda.synchronize()
    ctx.count(None, 32, device_ptr=dev.data_ptr(), nbytes=fa.nbytes)
    k3, c3 = ctx.table_host()
    assert np.array_equal(k1, k2) and np.array_equal(c1, c2)
    assert np.array_equal(k1, k3) and np.array_equal(c1, c3)


def test_skewed_buckets(ctx, oracle_mod):
    """Identical and near-identical k-mers overflow a bucket's capacity and
    go through the split path."""
    _p("case: skewed buckets")
    rng = np.random.default_rng(5)
    parts = []
    for i in range(3000):
        parts.append(b">a%d\n" % i + b"A" * 100 + b"\n")
    bases = np.frombuffer(b"ACGT", dtype=np.uint8)
    for i in range(2000):
        s = bases[rng.integers(0, 4, 100)].tobytes()
        parts.append(b">r%d\n" % i + s + b"\n")
    # low-complexity: (AC)^n with sparse substitutions -> many near-identical keys
    for i in range(4000):
        s = bytearray(b"AC" * 75)
        for j in rng.integers(0, 150, 2):
            s[j] = b"ACGT"[rng.integers(0, 4)]
        parts.append(b">l%d\n" % i + bytes(s) + b"\n")
    fa = b"".join(parts)
    for k in (32, 64, 21):
        _check_table(ctx, oracle_mod, fa, k)


def test_device_resident_input(ctx, oracle_mod):
    _p("case: device-resident input")
    import torch
    fa = read_fasta("syn_mid.fa.gz")
    t = torch.from_numpy(np.frombuffer(fa, dtype=np.uint8).copy()).to("cuda:0")
    torch.cuda.synchronize()
    ctx.count(None, 32, device_ptr=t.data_ptr(), nbytes=t.numel())
    keys, counts = ctx.table_host()
    ok, oc = oracle_mod.count(fa, 32)
    assert np.array_equal(keys, ok) and np.array_equal(counts, oc)


def test_small_config_full(ctx, oracle_mod, gen_mod):
    """1 Mbp genome, 100 bp reads, 30x, k=32, -t 21: table, cutoff and nodes."""
    _p("case: small config 1 Mbp 30x")
    fa = gen_mod.make_fasta(gen_mod.config_small())
    keys, counts = _check_table(ctx, oracle_mod, fa, 32)
    hist, mb, nk = ctx.cutoff(21)
    ohist, omb, okeep, _ = oracle_mod.cutoff(counts, 21)
    assert np.array_equal(hist, ohist) and mb == omb and nk == int(okeep.sum())
    kk, kc = ctx.kept_host()
    _check_nodes(ctx, oracle_mod, 32, kk, kc)


def test_repeat_family_k64(ctx, oracle_mod, gen_mod):
    """Scaled-down large-with-repeats workload at k=64 (two-word keys)."""
    _p("case: repeats k=64 (2 Mbp)")
    fa = gen_mod.make_fasta(gen_mod.config_large(scale=0.02))
    keys, counts = _check_table(ctx, oracle_mod, fa, 64)
    hist, mb, nk = ctx.cutoff(21)
    ohist, omb, okeep, _ = oracle_mod.cutoff(counts, 21)
    assert np.array_equal(hist, ohist) and mb == omb and nk == int(okeep.sum())
    kk, kc = ctx.kept_host()
    _check_nodes(ctx, oracle_mod, 64, kk, kc)


@pytest.mark.parametrize("case", ["small_k32", "repeats_k64", "skewed_k32"])
def test_persistent_partition_scatter(ctx, oracle_mod, gen_mod, case):
    """The persistent form of the further partition levels (each block walks
    many consecutive tiles, crossing bucket ends and empty buckets), forced
    at sizes where every block gets several tiles, against the oracle."""
    _p(f"case: persistent partition scatter {case}")
    if case == "small_k32":
        fa, k = gen_mod.make_fasta(gen_mod.config_small()), 32
    elif case == "repeats_k64":
        fa, k = gen_mod.make_fasta(gen_mod.config_large(scale=0.02)), 64
    else:
        # 3 in 8 reads one repeated k-mer, the rest random over A, C, G: one
        # huge bucket, and every level-1 bucket whose digit holds a T empty
        rng = np.random.default_rng(11)
        bases = np.frombuffer(b"ACG", dtype=np.uint8)
        parts = [b">a%d\n" % i + b"A" * 100 + b"\n" for i in range(60000)]
        parts += [b">r%d\n" % i + bases[rng.integers(0, 3, 100)].tobytes() + b"\n" for i in range(100000)]
        fa, k = b"".join(parts), 32
    try:
        ctx.set_tuning(ms_scatter=1, bucket_factor=1.5)
        _check_table(ctx, oracle_mod, fa, k)
        # more than two tiles (8192 one-word or 4096 two-word keys) for each
        # of the 512 blocks of a 256-CU device
        assert ctx.timing()["n_instances"] > 2 * 512 * (4096 if k > 32 else 8192)
    finally:
        ctx.set_tuning()


def test_bacterial_full_size_properties(ctx, gen_mod):
    """4.6 Mbp / 80x / k=32 at full size: size-independent properties (sum of
    counts = number of k-mer instances, strictly ascending keys, counts >= 1,
    kept table = filter of the table, node count bounds)."""
    _p("case: bacterial full size properties")
    spec = gen_mod.config_bacterial()
    fa = gen_mod.make_fasta(spec)
    R = spec.num_reads()
    ctx.count(fa, 32)
    keys, counts = ctx.table_host()
    assert int(counts.astype(np.int64).sum()) == R * (spec.read_len - 32 + 1)
    assert np.all(keys[1:, 0] > keys[:-1, 0]) and counts.min() >= 1
    hist, mb, nk = ctx.cutoff(21)
    h = np.bincount(counts[counts < 21], minlength=21)
    assert np.array_equal(hist, h.astype(np.uint64))
    assert mb == 1 + int(np.argmin(h[1:]))
    kk, kc = ctx.kept_host()
    sel = counts >= mb
    assert np.array_equal(kk, keys[sel]) and np.array_equal(kc, counts[sel])
    m = ctx.macro_nodes()
    nodes = ctx.nodes_host()
    assert nk <= m <= 2 * nk
    # every kept k-mer appears exactly once as a suffix and once as a prefix
    assert int((nodes["suf_cnt"] > 0).sum()) == nk and int((nodes["pre_cnt"] > 0).sum()) == nk
    assert int(nodes["suf_cnt"].astype(np.int64).sum()) == int(kc.astype(np.int64).sum())
    assert np.all(nodes["key_lo"][1:] > nodes["key_lo"][:-1])


NON_ACGT = (b">h\n" + b"ACGTNNacgtRYKMACGTTGCAAGGCTTAACCGGTTAA" * 6 + b"\n>h2\n"
            + b"nNnNACGTTTGGCCAAGGTTCCAAGGTTACGATCGATCGGATCAGCTAGC" * 4 + b"\n")


@pytest.mark.parametrize("name", ["syn_tiny.fa.gz", "syn_mid.fa.gz"])
def test_lmer_frequency(ctx, oracle_mod, name):
    _p(f"case: lmer frequency {name}")
    fa = read_fasta(name)
    assert np.array_equal(ctx.lmer_frequency(fa), oracle_mod.lmer_frequency(fa))
    assert np.array_equal(ctx.lmer_frequency(NON_ACGT), oracle_mod.lmer_frequency(NON_ACGT))


@pytest.mark.parametrize("k,P", [(32, 2), (32, 3), (21, 8), (64, 3), (45, 2)])
def test_rank_share(ctx, oracle_mod, k, P):
    """One rank's share (owner = hash31(min-frequency 8-mer) % P) equals the
    oracle's, and on ACGT-only reads the shares partition the whole table."""
    _p(f"case: rank share k={k} P={P}")
    for name in ("syn_mid.fa.gz", None):
        fa = read_fasta(name) if name else NON_ACGT
        total = 0
        for r in range(P):
            ctx.count_rank_share(fa, k, P, r)
            keys, counts = ctx.table_host()
            ok, oc = oracle_mod.count(fa, k, nranks=P, rank=r)
            assert np.array_equal(keys, ok) and np.array_equal(counts, oc), (name, r)
            total += keys.shape[0]
        ctx.count(fa, k)
        if name:
            assert total == ctx.table.n
        else:
            # with non-ACGT bytes the owner also depends on bases outside the
            # k-mer (the 3-bit code of the rolling 8-mer), so one k-mer value
            # can reach two ranks -- the oracle agrees rank by rank
            assert total >= ctx.table.n


def test_rank_share_non_acgt_vs_restatement(ctx):
    """Min-L-mer routing and L-mer frequencies on reads with non-ACGT bytes,
    against the window-by-window restatement of tests/rank_util.py (not the
    oracle)."""
    import rank_util as RU
    from test_rank_cpu import NON_ACGT as FA
    _p("case: rank share non-ACGT vs restatement")
    freq = RU.lmer_freq(FA)
    assert np.array_equal(ctx.lmer_frequency(FA), freq)
    for k, P in ((8, 3), (21, 3), (32, 2), (33, 5), (64, 3)):
        walk = RU.owner_walk(FA, k, P, freq)
        for r in range(P):
            ctx.count_rank_share(FA, k, P, r)
            keys, counts = ctx.table_host()
            assert RU.table_counter(keys, counts) == RU.share_counter(walk, r), (k, P, r)


@pytest.mark.parametrize("k,P", [(21, 3), (32, 2), (32, 8), (33, 3), (64, 5)])
def test_rank_share_by_key(ctx, oracle_mod, k, P):
    """The -p counters' owner rule (retrieve_proc_id of the whole k-mer,
    kmer_counting.cpp:1949-1950): each share is the one-rank table filtered
    by the closed-form owner (hash31 / uhash31 of tests/rank_util.py), and
    the shares partition the table."""
    import rank_util as RU
    _p(f"case: rank share by key k={k} P={P}")
    for fa in (read_fasta("syn_mid.fa.gz"), NON_ACGT):
        ok, oc = oracle_mod.count(fa, k)
        full = RU.table_counter(ok, oc)
        total = 0
        for r in range(P):
            ctx.count_rank_share_by_key(fa, k, P, r)
            keys, counts = ctx.table_host()
            got = RU.table_counter(keys, counts)
            assert got == {x: c for x, c in full.items() if RU.key_owner(x, k, P) == r}, r
            # ascending within the share
            vals = [int(x[0]) | ((int(x[1]) << 64) if keys.shape[1] == 2 else 0) for x in keys]
            assert vals == sorted(set(vals))
            total += len(got)
        assert total == len(full)


@pytest.mark.parametrize("name,k,t,P", [("syn_mid.fa.gz", 32, 21, 3), ("syn_rep.fa.gz", 64, 15, 4),
                                        ("syn_tiny.fa.gz", 33, 21, 2)])
def test_macro_nodes_rank_share(ctx, oracle_mod, name, k, t, P):
    """The nodes a rank holds at P ranks: node keys owned by retrieve_proc_id
    (hash31 for k <= 32, uhash31 above), each with its one-rank record."""
    import rank_util as RU
    _p(f"case: macro-node rank share {name} k={k} P={P}")
    from pakman_plus_amd import api
    fa = read_fasta(name)
    ok, oc = oracle_mod.count(fa, k)
    _, _, keep, _ = oracle_mod.cutoff(oc, t)
    want = oracle_mod.macro_nodes(ok[keep], oc[keep], k)
    wkeys = [int(x[0]) | ((int(x[1]) << 64) if k > 32 else 0) for x in want["keys"]]
    seen = 0
    for r in range(P):
        ctx.count(fa, k)
        ctx.cutoff(t)
        m = ctx.macro_nodes_rank_share(P, r)
        got = api.nodes_as_arrays(ctx.nodes_host(), k)
        sel = np.array([RU.key_owner(x, k, P) == r for x in wkeys], dtype=bool)
        assert m == int(sel.sum())
        for f in ("keys", "pre_cnt", "suf_cnt"):
            assert np.array_equal(got[f], want[f][sel]), (r, f)
        seen += m
    assert seen == len(wkeys)


@pytest.mark.parametrize("tuning", [dict(l1_chunks=1, l1_bits=8, bucket_factor=1.5),
                                    dict(l1_chunks=3, l1_bits=6, bucket_factor=5.0),
                                    dict(l1_chunks=8, l1_bits=4, bucket_factor=0.5),
                                    dict(l1_chunks=2, l1_bits=10, bucket_factor=2.2),
                                    # buckets far beyond a hash table: every
                                    # bucket overflows its first passes
                                    dict(l1_chunks=1, l1_bits=9, bucket_factor=24.0),
                                    # both forms of the further levels' scatter
                                    dict(ms_scatter=1), dict(ms_scatter=2, bucket_factor=1.5),
                                    # tiny buckets: three partition levels with the
                                    # built-in level-1 width; the second level's
                                    # counts folded into the level-1 histogram
                                    # (built-in) or counted in their own pass
                                    dict(bucket_factor=0.001), dict(bucket_factor=0.001, ms_fold=2),
                                    dict(bucket_factor=0.003, l1_bits=3),
                                    dict(bucket_factor=0.001, ms_scatter=2)])
def test_tuning_does_not_change_results(ctx, oracle_mod, tuning):
    """Launch-shape choices of the autotuner leave every output unchanged."""
    _p(f"case: tuning {tuning}")
    fa = read_fasta("syn_mid.fa.gz")
    try:
        ctx.set_tuning(**tuning)
        for k in (32, 64):
            keys, counts = _check_table(ctx, oracle_mod, fa, k)
            hist, mb, nk = ctx.cutoff(21)
            ohist, omb, okeep, _ = oracle_mod.cutoff(counts, 21)
            assert np.array_equal(hist, ohist) and mb == omb and nk == int(okeep.sum())
            kk, kc = ctx.kept_host()
            _check_nodes(ctx, oracle_mod, k, kk, kc)
    finally:
        ctx.set_tuning()


def test_tuning_out_of_range(ctx):
    from pakman_plus_amd import api
    for bad in (dict(l1_bits=11), dict(l1_chunks=65), dict(bucket_factor=65.0), dict(ms_scatter=3),
                dict(ms_scatter=-1)):
        with pytest.raises(api.PakHipError) as e:
            ctx.set_tuning(**bad)
        assert e.value.code == api.PK_EINVAL
    ctx.set_tuning()


@pytest.mark.parametrize("name,k,cov,t", [(n, k, c, t) for n, (c, ks, t) in SYNTH.items() for k in ks])
def test_initial_wiring_matches_reference(ctx, oracle_mod, name, k, cov, t):
    """pk_mn_wiring with the reference's prefix order reproduces the wiring the
    reference's own initiate_mnode_wiring / setup_wiring wrote (node dumps in
    tests/golden): entry counts, terminal (count, visit), every wire and every
    prefix's (first wire, number of wires)."""
    _p(f"case: initial wiring {name} k={k}")
    fa = read_fasta(f"syn_{name}.fa.gz")
    ref = read_ref(f"syn_{name}_k{k}.ref.txt.gz")
    ctx.count(fa, k)
    ctx.cutoff(t)
    kk, kc = ctx.kept_host()
    m = ctx.macro_nodes()
    nodes = ctx.nodes_host()
    order = oracle_mod.macro_nodes(kk, kc, k, ref_order=True)["pre_order"]
    w = ctx.mn_wiring(cov, order)
    assert m == len(ref.nodes)
    index = {(int(lo), int(hi)): i for i, (lo, hi) in enumerate(zip(nodes["key_lo"], nodes["key_hi"]))}
    multi = 0
    for node in ref.nodes:
        i = index[node.key]
        n_p, n_s = len(node.prefixes), len(node.suffixes)
        assert (w["n_prefix"][i], w["n_suffix"][i]) == (n_p, n_s)
        assert tuple(w["pre_terminal"][i]) == tuple(node.prefixes[-1][1:3])
        assert tuple(w["suf_terminal"][i]) == tuple(node.suffixes[-1][1:3])
        got = [tuple(int(x) for x in e) for e in w["wire"][i][:n_p + n_s + 1]]
        assert got == node.wiring, (node.k1mer, got, node.wiring)
        assert [tuple(int(x) for x in e) for e in w["prefix_begin"][i][:n_p]] == node.begin
        multi += int(n_p > 2 or n_s > 2)
    if name == "rep":
        assert multi > 0  # repeat families leave branching nodes


def test_count_file(ctx, oracle_mod, gen_mod, tmp_path):
    """pk_count_file (file -> pinned staging -> HBM) gives pk_count's table;
    bad paths fail with PK_EINVAL."""
    from pakman_plus_amd import api
    _p("case: count from a file")
    for name, fa in (("mid", read_fasta("syn_mid.fa.gz")),
                     ("small", bytes(gen_mod.make_fasta(gen_mod.config_small())))):  # 33 MB: staged
        f = tmp_path / f"{name}.fa"
        f.write_bytes(fa)
        for k in (32, 64):
            ctx.count_file(str(f), k)
            keys, counts = ctx.table_host()
            ctx.count(fa, k)
            k2, c2 = ctx.table_host()
            assert np.array_equal(keys, k2) and np.array_equal(counts, c2), (name, k)
    e = tmp_path / "empty.fa"
    e.write_bytes(b"")
    assert ctx.count_file(str(e), 32).n == 0
    for bad in (str(tmp_path / "missing.fa"), str(tmp_path)):
        with pytest.raises(api.PakHipError) as err:
            ctx.count_file(bad, 32)
        assert err.value.code == api.PK_EINVAL


def test_stream_ordered_alloc_mode(oracle_mod):
    """A context in the stream-ordered allocation mode (the server's slots)
    counts the same tables, also while its buffers grow call by call."""
    from pakman_plus_amd import api
    _p("case: stream-ordered allocation mode")
    c = api.Context(0)
    try:
        c.set_alloc_mode(True)
        for name, k in (("syn_tiny.fa.gz", 32), ("syn_mid.fa.gz", 64), ("syn_rep.fa.gz", 32)):
            fa = read_fasta(name)
            c.count(fa, k)
            keys, counts = c.table_host()
            ok, oc = oracle_mod.count(fa, k)
            assert np.array_equal(keys, ok) and np.array_equal(counts, oc), name
        c.set_alloc_mode(False)
        c.count(fa, k)
        assert c.table.n == ok.shape[0]
    finally:
        c.close()
