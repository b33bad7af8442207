"""The large-with-repeats workload at full size (100 Mbp genome, 5% from 50
repeat families at 2% divergence, 150 bp reads, 50x, 0.5% errors, k = 64:
3.3e7 reads, 2.9e9 k-mer instances, ~1e9 distinct two-word keys).

The oracle cannot count this input in test time, so the table, the cutoff
and the macro nodes are checked through properties that hold at any size:
  * sum of counts == reads * (150 - 64 + 1), counts >= 1;
  * keys strictly ascending as unsigned 128-bit values;
  * histogram == bincount of the counts below t, min_bucket == first argmin
    of hist[1..t-1], kept table == the entries with count >= min_bucket, in
    table order;
  * every kept k-mer is the suffix of exactly one node and the prefix of
    exactly one node (so the counts of either side sum to the kept counts),
    node keys strictly ascending;
  * a second count with the other partition-scatter form and without the
    level-2 histogram fold gives the same table (digest of keys and counts).
The same pipeline at 2 % of this size is compared with the oracle entry for
entry in test_gpu_parity.py (test_repeat_family_k64)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHUNK = 1 << 26  # entries per host-side check step


def _p(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ascending_u128(lo: np.ndarray, hi: np.ndarray) -> bool:
    """Strictly ascending (hi, lo) pairs, checked in chunks that overlap by one."""
    n = lo.shape[0]
    for s in range(0, max(n - 1, 0), CHUNK):
        e = min(n, s + CHUNK + 1)
        h0, h1 = hi[s:e - 1], hi[s + 1:e]
        l0, l1 = lo[s:e - 1], lo[s + 1:e]
        if not np.all((h1 > h0) | ((h1 == h0) & (l1 > l0))):
            return False
    return True


def _digest(keys: np.ndarray, counts: np.ndarray) -> tuple:
    """Order-sensitive digest of a table: wrapping sums of key words weighted
    by position and count, and the XOR of the words."""
    acc = [np.uint64(0)] * 4
    n = counts.shape[0]
    with np.errstate(over="ignore"):
        for s in range(0, n, CHUNK):
            e = min(n, s + CHUNK)
            idx = np.arange(s, e, dtype=np.uint64) + np.uint64(1)
            c = counts[s:e].astype(np.uint64)
            lo, hi = keys[s:e, 0], keys[s:e, 1]
            acc[0] = acc[0] + np.sum(lo * idx, dtype=np.uint64)
            acc[1] = acc[1] + np.sum(hi * c, dtype=np.uint64)
            acc[2] = acc[2] ^ np.bitwise_xor.reduce(lo ^ (hi << np.uint64(1)))
            acc[3] = acc[3] + np.sum(c * idx, dtype=np.uint64)
    return tuple(int(x) for x in acc)


@pytest.fixture(scope="module")
def large():
    from pakman_plus_amd import api, gen
    _p("case: large k=64 full size: generating 100 Mbp / 50x reads")
    spec = gen.config_large()
    fa = gen.make_fasta(spec)
    ctx = api.Context(0)
    yield ctx, spec, fa
    ctx.release_buffers()
    ctx.close()


@pytest.mark.timeout(900)
def test_large_full_size_properties(large):
    ctx, spec, fa = large
    k, t = 64, 21
    R = spec.num_reads()
    _p(f"case: large k=64 full size: count ({fa.nbytes} bytes, {R} reads)")
    ctx.count(fa, k)
    tm = ctx.timing()
    assert tm["n_reads"] == R and tm["n_instances"] == R * (spec.read_len - k + 1)
    keys, counts = ctx.table_host()
    U = counts.shape[0]
    _p(f"  table: {U} distinct, {tm['ms_levels']} further partition levels, {tm['n_split']} split buckets")
    assert int(counts.sum(dtype=np.int64)) == R * (spec.read_len - k + 1)
    assert int(counts.min()) >= 1
    lo = np.ascontiguousarray(keys[:, 0])
    hi = np.ascontiguousarray(keys[:, 1])
    assert _ascending_u128(lo, hi), "keys not strictly ascending"
    # keys hold 2k = 128 bits: no bit above the mask to check at k = 64
    _p("  cutoff")
    hist, mb, nk = ctx.cutoff(t)
    h = np.zeros(t, dtype=np.uint64)
    for s in range(0, U, CHUNK):
        c = counts[s:s + CHUNK]
        h += np.bincount(c[c < t], minlength=t).astype(np.uint64)
    assert np.array_equal(hist, h)
    assert mb == 1 + int(np.argmin(h[1:]))
    kk, kc = ctx.kept_host()
    assert kk.shape[0] == nk
    off = 0
    for s in range(0, U, CHUNK):
        c = counts[s:s + CHUNK]
        sel = c >= mb
        m = int(sel.sum())
        assert np.array_equal(kc[off:off + m], c[sel])
        assert np.array_equal(kk[off:off + m], keys[s:s + CHUNK][sel])
        off += m
    assert off == nk
    _p(f"  kept {nk} (min_bucket {mb}); macro nodes")
    m = ctx.macro_nodes()
    nodes = ctx.nodes_host()
    assert nk <= m <= 2 * nk
    assert int((nodes["suf_cnt"] > 0).sum()) == nk and int((nodes["pre_cnt"] > 0).sum()) == nk
    ks = int(kc.sum(dtype=np.int64))
    assert int(nodes["suf_cnt"].sum(dtype=np.int64)) == ks
    assert int(nodes["pre_cnt"].sum(dtype=np.int64)) == ks
    assert _ascending_u128(np.ascontiguousarray(nodes["key_lo"]), np.ascontiguousarray(nodes["key_hi"]))
    # (k-1)-mer keys hold 126 bits
    assert int(nodes["key_hi"].max()) >> 62 == 0
    del nodes, kk, kc
    d0 = _digest(keys, counts)
    del keys, counts, lo, hi
    _p(f"  {m} nodes; recount with one block per tile in the further partition levels, no fold")
    try:
        ctx.set_tuning(ms_scatter=2, ms_fold=2)
        ctx.count(fa, k)
        keys2, counts2 = ctx.table_host()
    finally:
        ctx.set_tuning()
    assert counts2.shape[0] == U
    assert _digest(keys2, counts2) == d0
    _p("  done")
