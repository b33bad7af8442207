"""ctypes wrapper of the CPU oracle (oracle/pk_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the pakman_plus_amd package.

Besides the wrapper, normalize() restates the input-format rewrite of
pk_count_format in Python; it is pinned by known answers and by round trips
(strict FASTA unchanged; FASTQ and wrapped multi-line forms of the golden
sets, with '\n' or '\r\n', rewrite to the original: tests/test_formats_cpu.py).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpk_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

PK_OK, PK_EINVAL, PK_ENOMEM, PK_EFORMAT, PK_EEMPTY = 0, -1, -2, -4, -5


def build(force: bool = False) -> str:
    """Compile pk_oracle.cpp with g++ into oracle/libpk_oracle.so."""
    src = os.path.join(HERE, "pk_oracle.cpp")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        tmp = LIB_PATH + ".tmp"
        subprocess.check_call(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-pthread", src, "-o", tmp])
        os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_ref() -> None:
    """Build oracle/_ref from the reference tree when it is present."""
    subprocess.check_call(["bash", os.path.join(HERE, "build_ref.sh")])


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, i32, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64
        pp = ctypes.POINTER(ctypes.c_void_p)
        L.pko_count.argtypes = [vp, sz, i32, i32, i32, u64, i32, pp]
        L.pko_bruteforce_count.argtypes = [vp, sz, i32, pp]
        L.pko_builder_count.argtypes = [vp, sz, i32, u64, i32, ctypes.c_char, ctypes.c_char, pp]
        for f in ("pko_count", "pko_bruteforce_count", "pko_builder_count", "pko_cutoff",
                  "pko_macro_nodes", "pko_lmer_frequency", "pko_lmer_owner", "pko_visit_value"):
            getattr(L, f).restype = i32
        L.pko_table_size.argtypes = [vp]
        L.pko_table_size.restype = sz
        L.pko_table_words.argtypes = [vp]
        L.pko_table_words.restype = i32
        L.pko_table_reads.argtypes = [vp]
        L.pko_table_reads.restype = u64
        L.pko_table_copy.argtypes = [vp, vp, vp]
        L.pko_table_copy.restype = None
        L.pko_table_free.argtypes = [vp]
        L.pko_table_free.restype = None
        L.pko_cutoff.argtypes = [vp, sz, i32, vp, ctypes.POINTER(i32), vp, ctypes.POINTER(sz)]
        L.pko_macro_nodes.argtypes = [vp, vp, sz, i32, i32, pp]
        L.pko_nodes_size.argtypes = [vp]
        L.pko_nodes_size.restype = sz
        L.pko_nodes_copy.argtypes = [vp, vp, vp, vp, vp, vp]
        L.pko_nodes_copy.restype = None
        L.pko_nodes_free.argtypes = [vp]
        L.pko_nodes_free.restype = None
        L.pko_lmer_frequency.argtypes = [vp, sz, vp]
        L.pko_lmer_owner.argtypes = [u64, i32]
        L.pko_visit_value.argtypes = [i32, i32]
        L.pko_pipeline_ranks.argtypes = [vp, sz, i32, i32, i32, vp, vp]
        L.pko_pipeline_ranks.restype = i32
        _LIB = L
    return _LIB


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed with code {code}")
        self.code = code


def _buf(data) -> tuple[np.ndarray, int]:
    a = np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray)) else np.asarray(data, dtype=np.uint8)
    a = np.ascontiguousarray(a)
    return a, a.nbytes


def _take_table(h):
    L = lib()
    n = L.pko_table_size(h)
    w = L.pko_table_words(h)
    reads = L.pko_table_reads(h)
    keys = np.empty((n, w), dtype=np.uint64)
    counts = np.empty(n, dtype=np.int32)
    L.pko_table_copy(h, keys.ctypes.data, counts.ctypes.data)
    L.pko_table_free(h)
    return keys, counts, reads


def count(fasta, k: int, nranks: int = 1, rank: int = 0, batch: int = 100_000_000,
          inplace: bool = False):
    """Sorted unique (keys[n, words], counts[n]) of the main path."""
    a, n = _buf(fasta)
    h = ctypes.c_void_p()
    rc = lib().pko_count(a.ctypes.data, n, k, nranks, rank, batch, int(inplace), ctypes.byref(h))
    if rc:
        raise OracleError(rc, "pko_count")
    keys, counts, _ = _take_table(h)
    return keys, counts


def bruteforce_count(fasta, k: int):
    a, n = _buf(fasta)
    h = ctypes.c_void_p()
    rc = lib().pko_bruteforce_count(a.ctypes.data, n, k, ctypes.byref(h))
    if rc:
        raise OracleError(rc, "pko_bruteforce_count")
    keys, counts, _ = _take_table(h)
    return keys, counts


def builder_count(fasta, klen: int, read_len: int = 100, kmer_bits: int = 64,
                  delimiter: bytes = b"\n", skip_symbol: bytes = b">"):
    """(reads, keys, counts) of Concurrent_Kmer_Partial_Counter::default_count_and_process."""
    a, n = _buf(fasta)
    h = ctypes.c_void_p()
    rc = lib().pko_builder_count(a.ctypes.data, n, klen, read_len, kmer_bits, delimiter, skip_symbol,
                                 ctypes.byref(h))
    if rc:
        raise OracleError(rc, "pko_builder_count")
    keys, counts, reads = _take_table(h)
    return int(reads), keys, counts


def cutoff(counts: np.ndarray, t: int):
    """(hist[t], min_bucket, keep_mask) -- raises OracleError(PK_EEMPTY) if nothing is kept."""
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    hist = np.zeros(max(t, 1), dtype=np.uint64)
    keep = np.zeros(counts.size, dtype=np.uint8)
    mb = ctypes.c_int()
    nk = ctypes.c_size_t()
    rc = lib().pko_cutoff(counts.ctypes.data, counts.size, t, hist.ctypes.data, ctypes.byref(mb),
                          keep.ctypes.data, ctypes.byref(nk))
    if rc and rc != PK_EEMPTY:
        raise OracleError(rc, "pko_cutoff")
    return hist, int(mb.value), keep.astype(bool), rc


def macro_nodes(keys: np.ndarray, counts: np.ndarray, k: int, ref_order: bool = False):
    """Dict of arrays: keys[m, words], pre_cnt[m,4], suf_cnt[m,4], pre_order[m,4], suf_order[m,4]."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    h = ctypes.c_void_p()
    rc = lib().pko_macro_nodes(keys.ctypes.data, counts.ctypes.data, counts.size, k, int(ref_order),
                               ctypes.byref(h))
    if rc:
        raise OracleError(rc, "pko_macro_nodes")
    L = lib()
    m = L.pko_nodes_size(h)
    w = 1 if k <= 32 else 2
    out = {
        "keys": np.empty((m, w), dtype=np.uint64),
        "pre_cnt": np.empty((m, 4), dtype=np.int32),
        "suf_cnt": np.empty((m, 4), dtype=np.int32),
        "pre_order": np.empty((m, 4), dtype=np.uint8),
        "suf_order": np.empty((m, 4), dtype=np.uint8),
    }
    L.pko_nodes_copy(h, out["keys"].ctypes.data, out["pre_cnt"].ctypes.data, out["suf_cnt"].ctypes.data,
                     out["pre_order"].ctypes.data, out["suf_order"].ctypes.data)
    L.pko_nodes_free(h)
    return out


def pipeline_ranks(fasta, k: int, t: int, nranks: int) -> dict:
    """The reference's count -> cutoff -> macro nodes at nranks ranks, one host
    thread per rank (the CPU baseline): distinct / kept / nodes / instances
    totals over the ranks, min_bucket, hist."""
    a, n = _buf(fasta)
    stats = np.zeros(5, dtype=np.uint64)
    hist = np.zeros(max(t, 1), dtype=np.uint64)
    rc = lib().pko_pipeline_ranks(a.ctypes.data, n, k, t, nranks, stats.ctypes.data, hist.ctypes.data)
    if rc and rc != PK_EEMPTY:
        raise OracleError(rc, "pko_pipeline_ranks")
    return {"distinct": int(stats[0]), "kept": int(stats[1]), "nodes": int(stats[2]),
            "instances": int(stats[3]), "min_bucket": int(stats[4]), "hist": hist, "rc": rc}


def lmer_frequency(fasta) -> np.ndarray:
    a, n = _buf(fasta)
    f = np.zeros(1 << 16, dtype=np.uint64)
    rc = lib().pko_lmer_frequency(a.ctypes.data, n, f.ctypes.data)
    if rc:
        raise OracleError(rc, "pko_lmer_frequency")
    return f


def lmer_owner(lmer: int, nranks: int) -> int:
    return int(lib().pko_lmer_owner(lmer, nranks))


def visit_value(count: int, coverage: int) -> int:
    return int(lib().pko_visit_value(count, coverage))


def ref_binary(k: int) -> str | None:
    p = os.path.join(REF_DIR, f"pakref_k{k}")
    return p if os.path.exists(p) else None


def normalize(data, fmt: str) -> bytes:
    """Input-format rewrite into strict single-line FASTA, restated line by
    line on the CPU from the format rules of pk_count_format (include/
    pakhip.h).  The reference has no counterpart: it reads only strict
    single-line FASTA (src/input_reading.cpp:69-312, src/kmer_counting.cpp:
    2920-2945).  fmt: "fasta" (multi-line) or "fastq" (four-line records).
    Raises ValueError where the device reports PK_EFORMAT."""
    data = bytes(data)
    if not data:
        return b""
    lines = data.split(b"\n")
    if lines[-1] == b"":
        lines.pop()  # the empty tail after a final '\n' is not a line
    lines = [ln[:-1] if ln.endswith(b"\r") else ln for ln in lines]
    out = []
    if fmt == "fastq":
        if len(lines) % 4:
            raise ValueError("FASTQ line count is not a multiple of 4")
        for j in range(0, len(lines), 4):
            h, seq, plus = lines[j], lines[j + 1], lines[j + 2]
            if not h.startswith(b"@") or not plus.startswith(b"+"):
                raise ValueError("malformed FASTQ record")
            out.append(b">" + h[1:] + b"\n" + seq + b"\n")
        return b"".join(out)
    if fmt != "fasta":
        raise ValueError(fmt)
    header, seq = None, []
    for ln in lines:
        if ln.startswith(b">"):
            if header is not None:
                out.append(header + b"\n" + b"".join(seq) + b"\n")
            header, seq = ln, []
        elif header is not None:
            seq.append(ln)
    if header is not None:
        out.append(header + b"\n" + b"".join(seq) + b"\n")
    return b"".join(out)
