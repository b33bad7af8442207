"""Builder-API counting on the device (pk_builder_count, the drop-in for
Concurrent_Kmer_Partial_Counter::default_count_and_process,
src/concurrent_kmer_counter.hpp:2018-2082) against the CPU oracle, and the
C++ adapter (include/pakhip_builders.hpp) driven like the reference's own
builder callers.  Bit-exact: reads counted, keys, counts."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

from builder_util import builder_closed_form, edge_inputs, windows_closed_form
from golden_util import read_fasta

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADAPTER = os.path.join(ROOT, "tests", "cpp", "_bin", "builder_adapter_test")
PARAMS = [(1, 100, 64), (5, 30, 64), (21, 100, 64), (32, 100, 64), (33, 100, 64),
          (32, 100, 128), (40, 60, 128), (64, 100, 128)]


def _p(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from pakman_plus_amd import api
    c = api.Context(0)
    yield c
    c.close()


def _check(ctx, oracle_mod, fa, klen, read_len, bits, delim=b"\n", skip=b">"):
    reads = ctx.builder_count(fa, klen, read_len, bits, delim, skip)
    keys, counts = ctx.table_host()
    oreads, ok, oc = oracle_mod.builder_count(fa, klen, read_len, bits, delim, skip)
    assert reads == oreads
    assert keys.shape[0] == ok.shape[0]
    if ok.shape[0]:
        assert np.array_equal(keys, ok) and np.array_equal(counts, oc)
    return reads, keys, counts


@pytest.mark.parametrize("name", sorted(edge_inputs()))
@pytest.mark.parametrize("klen,read_len,bits", PARAMS)
def test_builder_edge_inputs(ctx, oracle_mod, name, klen, read_len, bits):
    _p(f"case: builder edge {name} klen={klen} read_len={read_len} bits={bits}")
    _check(ctx, oracle_mod, edge_inputs()[name], klen, read_len, bits)


def test_builder_custom_symbols(ctx, oracle_mod):
    _p("case: builder custom delimiter/skip")
    fa = edge_inputs()["ragged"].replace(b"\n", b"|").replace(b">", b"@")
    reads, _, _ = _check(ctx, oracle_mod, fa, 21, 100, 64, b"|", b"@")
    assert reads > 0
    from pakman_plus_amd import api
    with pytest.raises(api.PakHipError) as e:
        ctx.builder_count(fa, 21, 100, 64, b"@", b"@")
    assert e.value.code == api.PK_EINVAL


def test_builder_reference_unit_test(ctx, oracle_mod):
    """src/concurrent_kmer_counter_test.cpp:100-113: sum(freq) == reads*(100-k+1)."""
    _p("case: builder reference unit test")
    fa = read_fasta("ref_test_two_reads.fa")
    reads, keys, counts = _check(ctx, oracle_mod, fa, 32, 100, 64)
    assert reads == 2 and int(counts.sum()) == reads * (100 - 32 + 1)


@pytest.mark.parametrize("klen,bits", [(32, 64), (21, 64), (64, 128), (31, 128)])
def test_builder_synthetic_sets(ctx, oracle_mod, klen, bits):
    _p(f"case: builder syn_mid klen={klen} bits={bits}")
    fa = read_fasta("syn_mid.fa.gz")
    _check(ctx, oracle_mod, fa, klen, 100, bits)


def test_builder_ragged_reads_larger(ctx, oracle_mod, gen_mod):
    """Reads of 75/150/250 bp with read_len 100: short reads run into the
    next header, long ones are cut (the reference's fixed stride)."""
    _p("case: builder ragged read lengths")
    parts = []
    for i, L in enumerate((75, 150, 250)):
        spec = gen_mod.ReadSetSpec(genome_len=30_000, read_len=L, coverage=20.0, err=0.005,
                                   genome_seed=300 + i, read_seed=400 + i)
        parts.append(gen_mod.make_fasta(spec).tobytes())
    fa = b"".join(parts)
    for klen, bits in ((32, 64), (27, 64), (50, 128)):
        _check(ctx, oracle_mod, fa, klen, 100, bits)


def test_builder_table_feeds_cutoff_and_macro_nodes(ctx, oracle_mod):
    """klen 32 on 64-bit words over ACGT-only reads (150 bp, read_len 150): the
    builder table equals the main path's, and cutoff / macro nodes apply to it."""
    _p("case: builder table -> cutoff -> macro nodes")
    from pakman_plus_amd import api
    fa = read_fasta("syn_mid.fa.gz")
    ctx.count(fa, 32)
    mk, mc = ctx.table_host()
    ctx.builder_count(fa, 32, 150, 64)
    bk, bc = ctx.table_host()
    assert np.array_equal(mk, bk) and np.array_equal(mc, bc)
    hist, mb, nk = ctx.cutoff(21)
    ohist, omb, okeep, _ = oracle_mod.cutoff(bc, 21)
    assert np.array_equal(hist, ohist) and mb == omb and nk == int(okeep.sum())
    kk, kc = ctx.kept_host()
    m = ctx.macro_nodes()
    want = oracle_mod.macro_nodes(kk, kc, 32)
    got = api.nodes_as_arrays(ctx.nodes_host(), 32)
    assert m == want["keys"].shape[0]
    for f in ("keys", "pre_cnt", "suf_cnt"):
        assert np.array_equal(got[f], want[f]), f
    # klen 21 keeps 32 bases per word: no macro nodes on such a table
    ctx.builder_count(fa, 21, 100, 64)
    ctx.cutoff(21)
    with pytest.raises(api.PakHipError) as e:
        ctx.macro_nodes()
    assert e.value.code == api.PK_ESTATE


def _digest(keys, counts):
    x = 0
    M = (1 << 64) - 1
    for i in range(keys.shape[0]):
        w = int(keys[i, 0]) | ((int(keys[i, 1]) << 64) if keys.shape[1] == 2 else 0)
        x ^= (((w ^ (w >> 64)) & M) * 0x9E3779B97F4A7C15 + int(counts[i])) & M
    return x


def _run_adapter(*args):
    out = subprocess.run([ADAPTER, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    line = out.stdout.strip().splitlines()[-1]
    return {k: int(v) for k, v in (f.split("=") for f in line.split())}


def test_cpp_adapter_reference_shape(oracle_mod, tmp_path):
    """Hip_Kmer_Partial_Counter driven like default_kmer_table_builder.cpp:68-81."""
    _p("case: C++ adapter")
    r = _run_adapter()
    assert r["reads"] == 2 and r["freq"] == 138 and r["calls"] == 1
    fa = read_fasta("syn_mid.fa.gz")
    f = tmp_path / "mid.fa"
    f.write_bytes(fa)
    for klen, bits in ((32, 64), (64, 128)):
        oreads, ok, oc = oracle_mod.builder_count(fa, klen, 100, bits)
        r = _run_adapter(f, klen, 100, bits, 2 ** 64 - 1)
        assert r["reads"] == oreads and r["unique"] == ok.shape[0]
        assert r["freq"] == int(oc.astype(np.int64).sum()) and r["calls"] == 1
        assert r["xor"] == _digest(ok, oc)
    # slices of max_kmer_count pairs
    r = _run_adapter(f, 32, 100, 64, 1000)
    oreads, ok, oc = oracle_mod.builder_count(fa, 32, 100, 64)
    assert r["calls"] == -(-ok.shape[0] // 1000) and r["unique"] == ok.shape[0]
    assert r["xor"] == _digest(ok, oc)


def test_cpp_adapter_omp_v3(oracle_mod, tmp_path):
    """concurrent_count_and_process_omp_v3, as the reference's unit test calls
    it (src/concurrent_kmer_counter_test.cpp:88-111): same reads and table as
    the default counter, one processor call whatever max_kmer_count is, and
    the range is not consumed (two calls agree)."""
    _p("case: C++ adapter omp_v3")
    r = _run_adapter("v3")
    assert r["reads"] == 2 and r["freq"] == 138 and r["calls"] == 1
    fa = read_fasta("syn_mid.fa.gz")
    f = tmp_path / "mid.fa"
    f.write_bytes(fa)
    for klen, bits in ((32, 64), (64, 128)):
        oreads, ok, oc = oracle_mod.builder_count(fa, klen, 100, bits)
        r = _run_adapter("v3", f, klen, 100, bits, 1000)
        assert r["reads"] == oreads and r["unique"] == ok.shape[0] and r["calls"] == 1
        assert r["xor"] == _digest(ok, oc)


@pytest.mark.parametrize("name,k,threshold,cov", [("syn_mid.fa.gz", 32, 20, 100), ("syn_rep.fa.gz", 32, 14, 30),
                                                  ("syn_mid.fa.gz", 64, 10, 50)])
def test_pakgraph_nodes(oracle_mod, tmp_path, name, k, threshold, cov):
    """The PaK-graph builder's nodes (count > threshold, sorted entries,
    balanced terminals, wires) from the device records through the C++
    adapter, against the restatement in tests/pakgraph_util.py."""
    _p(f"case: pakgraph {name} k={k} threshold={threshold}")
    from pakgraph_util import pakgraph_lines
    fa = read_fasta(name)
    f = tmp_path / "in.fa"
    f.write_bytes(fa)
    out = subprocess.run([ADAPTER, "pakgraph", str(f), str(k), str(threshold), str(cov)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr
    got = out.stdout.strip().splitlines()
    want = pakgraph_lines(oracle_mod, fa, k, threshold, cov)
    assert len(got) == len(want) > 0
    assert got == want


def test_cutoff_threshold(ctx, oracle_mod):
    _p("case: threshold cutoff")
    fa = read_fasta("syn_mid.fa.gz")
    keys, counts = _check(ctx, oracle_mod, fa, 32, 150, 64)[1:]
    for thr in (0, 1, 20, 10 ** 6):
        nk = ctx.cutoff_threshold(thr)
        kk, kc = ctx.kept_host()
        sel = counts > thr  # pakman_graph.hpp:603 skips count <= threshold
        assert nk == int(sel.sum()) and np.array_equal(kk, keys[sel]) and np.array_equal(kc, counts[sel])


# -- the counters concurrent_kmer_table_builder.cpp:97-116 calls ---------------

WINDOW_PARAMS = [(4, 10), (11, 30), (21, 100), (2, 160)]


@pytest.mark.parametrize("name", sorted(edge_inputs()))
@pytest.mark.parametrize("klen,read_len", WINDOW_PARAMS)
@pytest.mark.parametrize("mode", ["v0", "v1"])
def test_builder_windows_edge_inputs(ctx, name, klen, read_len, mode):
    """pk_builder_count_windows against the strncmp-window restatement."""
    from pakman_plus_amd import api
    _p(f"case: builder windows {name} klen={klen} read_len={read_len} {mode}")
    fa = edge_inputs()[name]
    try:
        want = windows_closed_form(fa, klen, read_len, mode)
    except ValueError:
        with pytest.raises(api.PakHipError) as e:
            ctx.builder_count_windows(fa, klen, read_len, mode)
        assert e.value.code == api.PK_EFORMAT
        return
    try:
        got = ctx.builder_count_windows(fa, klen, read_len, mode)
    except api.PakHipError as e:
        # more byte values than a 128-bit key holds at this klen
        assert e.code == api.PK_EINVAL and "128-bit" in str(e)
        assert len(set(fa)) > 2 ** (128 // klen)
        return
    assert got == want


@pytest.mark.parametrize("klen,read_len", [(32, 100), (32, 150), (64, 150), (21, 100)])
def test_builder_windows_read_sets(ctx, klen, read_len):
    _p(f"case: builder windows syn_mid klen={klen} read_len={read_len}")
    fa = read_fasta("syn_mid.fa.gz")
    for mode in ("v0", "v1"):
        assert ctx.builder_count_windows(fa, klen, read_len, mode) == windows_closed_form(fa, klen, read_len, mode)


def test_builder_windows_nul_bytes(ctx):
    _p("case: builder windows with NUL bytes")
    fa = b">a\nAC\x00GTACGT\n>b\nAC\x00TTACGT\n>c\n\x00\x00ACGTACG\n"
    for klen in (2, 4, 7):
        assert ctx.builder_count_windows(fa, klen, 9, "v0") == windows_closed_form(fa, klen, 9, "v0")


def _run_mode(mode, path, klen, read_len, bits):
    out = subprocess.run([ADAPTER, mode, str(path), str(klen), str(read_len), str(bits)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    r = {}
    for ln in out.stdout.splitlines():
        if ln.startswith("unique kmers="):
            r["unique"] = int(ln.split("=")[1])
        elif ln.startswith("concurrent_vector size="):
            r["size"] = int(ln.split("=")[1])
        elif ln.startswith("reads="):
            r.update({k: int(v) for k, v in (f.split("=") for f in ln.split())})
    return r


@pytest.mark.parametrize("klen,bits", [(32, 64), (64, 128)])
def test_cpp_adapter_table_builder_modes(tmp_path, klen, bits):
    """The three counters of concurrent_kmer_table_builder.cpp:97-116 through
    the C++ adapter: reads returned, processor not called, the reference's
    "unique kmers" / "concurrent_vector size" figures at one thread."""
    _p(f"case: C++ adapter table-builder modes klen={klen}")
    fa = read_fasta("syn_mid.fa.gz")
    f = tmp_path / "mid.fa"
    f.write_bytes(fa)
    reads, counts = builder_closed_form(fa, klen, 100, bits)
    r = _run_mode("explicit", f, klen, 100, bits)
    assert r == {"unique": len(counts), "size": sum(counts.values()), "reads": reads, "calls": 0}
    for mode in ("v0", "v1"):
        wr, ww, wd = windows_closed_form(fa, klen, 100, mode)
        r = _run_mode("implicit_" + mode, f, klen, 100, bits)
        assert r == {"unique": wd, "size": ww, "reads": wr, "calls": 0}
    # an empty range: no k-mer, the reference's one-thread figure is 1
    e = tmp_path / "empty.fa"
    e.write_bytes(b"")
    assert _run_mode("explicit", e, klen, 100, bits) == {"unique": 1, "size": 0, "reads": 0, "calls": 0}
    assert _run_mode("implicit_v1", e, klen, 100, bits) == {"unique": 1, "size": 0, "reads": 0, "calls": 0}
