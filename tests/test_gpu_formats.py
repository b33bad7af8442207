"""pk_count_format / pk_normalize (multi-line FASTA, FASTQ, '\\r\\n' line
ends) through the C ABI: the device rewrite equals the CPU restatement byte
for byte, and the table equals pk_count's table of the strict single-line
FASTA of the same reads."""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

from format_util import to_fastq, to_multiline
from golden_util import read_fasta
from test_formats_cpu import KNOWN

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


@pytest.mark.parametrize("fmt,data,want", KNOWN)
def test_device_rewrite_known_answers(ctx, fmt, data, want):
    _p(f"case: rewrite {fmt} len={len(data)}")
    assert ctx.normalize(data, fmt) == want


def test_device_rewrite_random_lines(ctx, oracle_mod):
    """Random lines built from '>', '@', '+', '\\r' and bases, any layout."""
    rng = np.random.default_rng(3)
    alphabet = np.frombuffer(b">>@+\r\rACGTN", dtype=np.uint8)
    for trial in range(20):
        nlines = int(rng.integers(1, 400))
        lines = [rng.choice(alphabet, size=int(rng.integers(0, 90))).tobytes() for _ in range(nlines)]
        data = b"\n".join(lines) + (b"\n" if trial % 2 else b"")
        _p(f"case: random fasta trial {trial} len={len(data)}")
        assert ctx.normalize(data, "fasta") == oracle_mod.normalize(data, "fasta")


@pytest.mark.parametrize("name", ["syn_tiny.fa.gz", "syn_mid.fa.gz"])
def test_formats_count_like_strict_fasta(ctx, oracle_mod, name):
    from pakman_plus_amd import api
    fa = read_fasta(name)
    variants = {"fastq": to_fastq(fa), "fastq-crlf": to_fastq(fa, b"\r\n"),
                "fasta-60": to_multiline(fa, 60), "fasta-7-crlf": to_multiline(fa, 7, b"\r\n")}
    for k in (21, 32, 64):
        ctx.count(fa, k)
        want_k, want_c = ctx.table_host()
        ok, oc = oracle_mod.count(fa, k)
        assert np.array_equal(want_k, ok) and np.array_equal(want_c, oc)
        for vname, data in variants.items():
            _p(f"case: {name} {vname} k={k}")
            fmt = "fastq" if vname.startswith("fastq") else "fasta"
            assert ctx.normalize(data, fmt) == fa
            ctx.count_format(data, k, fmt)
            keys, counts = ctx.table_host()
            assert np.array_equal(keys, want_k) and np.array_equal(counts, want_c)
            assert ctx.timing()["normalize_ms"] > 0
    # strict passes through unchanged
    ctx.count_format(fa, 32, "fasta-strict")
    assert ctx.table.n == oracle_mod.count(fa, 32)[0].shape[0]
    assert api.FORMATS["fasta-strict"] == 0


def test_format_errors(ctx):
    from pakman_plus_amd import api
    for data in (b"@r\nAC\n+\n", b">r\nAC\n+\nII\n", b"@r\nAC\n-\nII\n"):
        _p(f"case: fastq error {data!r}")
        with pytest.raises(api.PakHipError) as e:
            ctx.count_format(data, 8, "fastq")
        assert e.value.code == api.PK_EFORMAT
    t = api._Table()
    data = b">a\nACGT\n"
    with pytest.raises(api.PakHipError) as e:
        ctx._check(ctx._lib.pk_count_format(ctx._h, data, len(data), 8, 7, 0, ctypes.byref(t)))
    assert e.value.code == api.PK_EINVAL


def test_fastq_bacterial_size(ctx, gen_mod):
    """The bacterial workload's reads as FASTQ with '\\r\\n' line ends count to
    the same table as its strict FASTA."""
    _p("case: bacterial fastq crlf")
    fa = gen_mod.make_fasta(gen_mod.config_bacterial()).tobytes()
    fq = to_fastq(fa, b"\r\n")
    ctx.count(fa, 32)
    k1, c1 = ctx.table_host()
    ctx.count_format(fq, 32, "fastq")
    k2, c2 = ctx.table_host()
    assert np.array_equal(k1, k2) and np.array_equal(c1, c2)


def test_device_resident_formats(ctx, oracle_mod):
    """Device-resident FASTQ / multi-line input (data_on_device) gives the
    strict FASTA's table, and leaves the caller's buffer unchanged."""
    import torch
    fa = read_fasta("syn_mid.fa.gz")
    ok, oc = oracle_mod.count(fa, 32)
    for fmt, data in (("fastq", to_fastq(fa, b"\r\n")), ("fasta", to_multiline(fa, 61))):
        _p(f"case: device-resident {fmt}")
        host = np.frombuffer(data, dtype=np.uint8)
        dev = torch.from_numpy(host.copy()).to("cuda:0")
        torch.cuda.synchronize()
        ctx.count_format(None, 32, fmt, device_ptr=dev.data_ptr(), nbytes=len(data))
        keys, counts = ctx.table_host()
        assert np.array_equal(keys, ok) and np.array_equal(counts, oc)
        assert np.array_equal(dev.cpu().numpy(), host)
