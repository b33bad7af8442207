"""Host-side mirror of the reference's counting / cutoff / macro-node interface
over the C ABI of libpakhip.so (include/pakhip.h).

Reference entry points mirrored (src/ paths of vkundeti/pakman_plus):
  perform_kmer_counting(read_data, length)      kmer_counting.cpp:3351
  construct_kmer_histogram(...)                 kmer_histogram.cpp:72
  construct_macro_nodes(MN_map, kmers)          mn_construct.cpp:212
  begin_mnode_construction(MN_map)              mn_construct.cpp:479
Errors are raised as PakHipError carrying the PK_* code; the reference's
aborts map to PK_EFORMAT (bad FASTA) and PK_EEMPTY (nothing kept).

The HIP library is the only implementation: if libpakhip.so is missing or no
GPU is present, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

PK_OK, PK_EINVAL, PK_ENOMEM, PK_EHIP, PK_EFORMAT, PK_EEMPTY, PK_ESTATE, PK_EAGAIN = 0, -1, -2, -3, -4, -5, -6, -7

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libpakhip.so")


class PakHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pakhip error {code}: {msg}")
        self.code = code


class _Table(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("words", ctypes.c_int32), ("k", ctypes.c_int32),
                ("keys", ctypes.c_void_p), ("counts", ctypes.c_void_p)]


class _MnTable(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("k", ctypes.c_int32), ("nodes", ctypes.c_void_p)]


class _Timing(ctypes.Structure):
    _fields_ = [("parse_ms", ctypes.c_float), ("hist_ms", ctypes.c_float),
                ("scatter_ms", ctypes.c_float), ("sort_ms", ctypes.c_float),
                ("split_ms", ctypes.c_float), ("compact_ms", ctypes.c_float),
                ("cutoff_ms", ctypes.c_float), ("mn_ms", ctypes.c_float),
                ("n_instances", ctypes.c_uint64), ("n_reads", ctypes.c_uint64),
                ("bucket_bits", ctypes.c_int32), ("n_split", ctypes.c_int32),
                ("l1_scatter_ms", ctypes.c_float), ("normalize_ms", ctypes.c_float),
                ("ms_hist_ms", ctypes.c_float), ("ms_scatter_ms", ctypes.c_float),
                ("ms_levels", ctypes.c_int32), ("ms_scatter_form", ctypes.c_int32),
                ("ms_level_hist_ms", ctypes.c_float * 4), ("ms_level_scatter_ms", ctypes.c_float * 4),
                ("ms_level_form", ctypes.c_int32 * 4)]


class Tuning(ctypes.Structure):
    """Launch-shape choices of the counting kernels (pk_tuning); 0 = built-in."""
    _fields_ = [("l1_chunks", ctypes.c_int32), ("l1_bits", ctypes.c_int32),
                ("bucket_factor", ctypes.c_float), ("ms_scatter", ctypes.c_int32),
                ("ms_fold", ctypes.c_int32)]


class _TuneEntry(ctypes.Structure):
    _fields_ = [("max_bytes", ctypes.c_uint64), ("key_words", ctypes.c_int32), ("tuning", Tuning)]


class _Job(ctypes.Structure):
    _fields_ = [("fasta", ctypes.c_void_p), ("nbytes", ctypes.c_uint64), ("k", ctypes.c_int32),
                ("t", ctypes.c_int32), ("fasta_on_device", ctypes.c_int32),
                ("want_nodes", ctypes.c_int32)]


class _JobResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("min_bucket", ctypes.c_int32),
                ("n_reads", ctypes.c_uint64), ("n_instances", ctypes.c_uint64),
                ("table", _Table), ("kept", _Table), ("nodes", _MnTable),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double),
                ("t_end", ctypes.c_double), ("est_bytes", ctypes.c_uint64),
                ("message", ctypes.c_char * 160)]


MNWIRE_DTYPE = np.dtype([("n_prefix", "<i4"), ("n_suffix", "<i4"), ("n_wire", "<i4"),
                         ("pre_terminal", "<i4", (2,)), ("suf_terminal", "<i4", (2,)),
                         ("wire", "<i4", (11, 3)), ("prefix_begin", "<i4", (5, 2))])

MNODE_DTYPE = np.dtype([("key_lo", "<u8"), ("key_hi", "<u8"), ("pre_cnt", "<i4", (4,)),
                        ("suf_cnt", "<i4", (4,))])

# every symbol declared in include/pakhip.h
EXPORTS = ("pk_ctx_create", "pk_ctx_destroy", "pk_last_error", "pk_ctx_stream",
           "pk_ctx_device_bytes", "pk_ctx_set_tuning", "pk_ctx_release_buffers", "pk_ctx_set_alloc_mode",
           "pk_count",
           "pk_count_noalloc", "pk_count_file",
           "pk_count_format", "pk_normalize", "pk_download_normalized",
           "pk_count_rank_share", "pk_count_rank_share_by_key", "pk_macro_nodes_rank_share",
           "pk_lmer_frequency", "pk_builder_count", "pk_builder_count_windows",
           "pk_cutoff", "pk_cutoff_threshold", "pk_kept_table", "pk_macro_nodes", "pk_mn_wiring", "pk_download_table", "pk_download_kept",
           "pk_download_nodes", "pk_last_timing", "pk_server_create", "pk_server_destroy",
           "pk_server_now", "pk_server_submit", "pk_server_submit_batch", "pk_server_wait",
           "pk_server_poll", "pk_server_download", "pk_server_set_tuning", "pk_server_release",
           "pk_server_last_error")

_LIB = None

# input formats of pk_count_format (PK_FMT_* in include/pakhip.h)
FORMATS = {"fasta-strict": 0, "fasta": 1, "fastq": 2}


def load_library():
    """Load libpakhip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    # torch ships its own HIP runtime under the same soname; when torch is
    # present it is loaded first so the process holds a single HIP runtime
    # that both torch and libpakhip bind to.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not found: build it with `python -m pakman_plus_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64
    L.pk_ctx_create.argtypes = [ctypes.POINTER(vp), i32, u64]
    L.pk_ctx_create.restype = i32
    L.pk_ctx_destroy.argtypes = [vp]
    L.pk_ctx_destroy.restype = None
    L.pk_last_error.argtypes = [vp]
    L.pk_last_error.restype = ctypes.c_char_p
    L.pk_ctx_stream.argtypes = [vp]
    L.pk_ctx_stream.restype = vp
    L.pk_ctx_device_bytes.argtypes = [vp]
    L.pk_ctx_device_bytes.restype = u64
    L.pk_count.argtypes = [vp, vp, u64, i32, i32, ctypes.POINTER(_Table)]
    L.pk_count.restype = i32
    L.pk_count_file.argtypes = [vp, ctypes.c_char_p, i32, ctypes.POINTER(_Table)]
    L.pk_count_file.restype = i32
    L.pk_count_noalloc.argtypes = [vp, vp, u64, i32, vp, vp, u64, ctypes.POINTER(u64)]
    L.pk_count_noalloc.restype = i32
    L.pk_count_format.argtypes = [vp, vp, u64, i32, i32, i32, ctypes.POINTER(_Table)]
    L.pk_count_format.restype = i32
    L.pk_normalize.argtypes = [vp, vp, u64, i32, i32, ctypes.POINTER(u64)]
    L.pk_normalize.restype = i32
    L.pk_download_normalized.argtypes = [vp, vp]
    L.pk_download_normalized.restype = i32
    L.pk_count_rank_share.argtypes = [vp, vp, u64, i32, i32, i32, i32, ctypes.POINTER(_Table)]
    L.pk_count_rank_share.restype = i32
    L.pk_count_rank_share_by_key.argtypes = [vp, vp, u64, i32, i32, i32, i32, ctypes.POINTER(_Table)]
    L.pk_count_rank_share_by_key.restype = i32
    L.pk_macro_nodes_rank_share.argtypes = [vp, i32, i32, ctypes.POINTER(_MnTable)]
    L.pk_macro_nodes_rank_share.restype = i32
    L.pk_lmer_frequency.argtypes = [vp, vp, u64, i32, vp]
    L.pk_lmer_frequency.restype = i32
    L.pk_builder_count.argtypes = [vp, vp, u64, i32, u64, i32, ctypes.c_char, ctypes.c_char, i32,
                                   ctypes.POINTER(_Table), ctypes.POINTER(u64)]
    L.pk_builder_count.restype = i32
    L.pk_builder_count_windows.argtypes = [vp, vp, u64, i32, u64, ctypes.c_char, ctypes.c_char, i32, i32,
                                           ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64)]
    L.pk_builder_count_windows.restype = i32
    L.pk_cutoff.argtypes = [vp, i32, vp, ctypes.POINTER(i32), ctypes.POINTER(u64)]
    L.pk_cutoff.restype = i32
    L.pk_cutoff_threshold.argtypes = [vp, u64, ctypes.POINTER(u64)]
    L.pk_cutoff_threshold.restype = i32
    L.pk_kept_table.argtypes = [vp, ctypes.POINTER(_Table)]
    L.pk_kept_table.restype = i32
    L.pk_macro_nodes.argtypes = [vp, ctypes.POINTER(_MnTable)]
    L.pk_macro_nodes.restype = i32
    L.pk_mn_wiring.argtypes = [vp, i32, vp, vp]
    L.pk_mn_wiring.restype = i32
    L.pk_download_table.argtypes = [vp, vp, vp]
    L.pk_download_table.restype = i32
    L.pk_download_kept.argtypes = [vp, vp, vp]
    L.pk_download_kept.restype = i32
    L.pk_download_nodes.argtypes = [vp, vp]
    L.pk_download_nodes.restype = i32
    L.pk_last_timing.argtypes = [vp, ctypes.POINTER(_Timing)]
    L.pk_last_timing.restype = i32
    L.pk_ctx_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.pk_ctx_set_tuning.restype = i32
    L.pk_server_set_tuning.argtypes = [vp, ctypes.POINTER(_TuneEntry), i32]
    L.pk_server_set_tuning.restype = i32
    L.pk_ctx_set_alloc_mode.argtypes = [vp, i32]
    L.pk_ctx_set_alloc_mode.restype = i32
    L.pk_ctx_release_buffers.argtypes = [vp]
    L.pk_ctx_release_buffers.restype = i32
    L.pk_server_create.argtypes = [ctypes.POINTER(vp), i32, i32, u64]
    L.pk_server_create.restype = i32
    L.pk_server_destroy.argtypes = [vp]
    L.pk_server_destroy.restype = None
    L.pk_server_now.argtypes = [vp]
    L.pk_server_now.restype = ctypes.c_double
    L.pk_server_submit.argtypes = [vp, ctypes.POINTER(_Job), ctypes.POINTER(u64)]
    L.pk_server_submit.restype = i32
    L.pk_server_submit_batch.argtypes = [vp, ctypes.POINTER(_Job), i32, ctypes.POINTER(u64)]
    L.pk_server_submit_batch.restype = i32
    L.pk_server_wait.argtypes = [vp, u64, ctypes.POINTER(_JobResult)]
    L.pk_server_wait.restype = i32
    L.pk_server_poll.argtypes = [vp, u64, ctypes.POINTER(_JobResult)]
    L.pk_server_poll.restype = i32
    L.pk_server_download.argtypes = [vp, u64, vp, vp, vp, vp, vp, vp]
    L.pk_server_download.restype = i32
    L.pk_server_release.argtypes = [vp, u64]
    L.pk_server_release.restype = i32
    L.pk_server_last_error.argtypes = [vp]
    L.pk_server_last_error.restype = ctypes.c_char_p
    _LIB = L
    return L


def _as_u8(fasta) -> np.ndarray:
    if isinstance(fasta, (bytes, bytearray)):
        return np.frombuffer(fasta, dtype=np.uint8)
    return np.ascontiguousarray(fasta, dtype=np.uint8)


@dataclass
class TableView:
    """Device-resident sorted (k-mer, count) table (pointers owned by the Context)."""
    n: int
    words: int
    k: int
    keys_ptr: int
    counts_ptr: int


class Context:
    """One device context (HIP stream + device memory) of libpakhip."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self._lib = load_library()
        h = ctypes.c_void_p()
        rc = self._lib.pk_ctx_create(ctypes.byref(h), device, hbm_budget_bytes)
        if rc:
            raise PakHipError(rc, "pk_ctx_create failed (no usable HIP device?)")
        self._h = h
        self.k = None
        self.table: TableView | None = None
        self.n_kept = 0
        self.n_nodes = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pk_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc:
            raise PakHipError(rc, self._lib.pk_last_error(self._h).decode(errors="replace"))

    @property
    def stream(self) -> int:
        return self._lib.pk_ctx_stream(self._h)

    @property
    def device_bytes(self) -> int:
        return int(self._lib.pk_ctx_device_bytes(self._h))

    def set_alloc_mode(self, stream_ordered: bool):
        """Buffers from the device's stream-ordered pool (True) or hipMalloc
        (False, the default); changing the mode frees the buffers."""
        self._check(self._lib.pk_ctx_set_alloc_mode(self._h, 1 if stream_ordered else 0))

    def release_buffers(self):
        """Free the context's device buffers (they grow back on the next call)."""
        self._check(self._lib.pk_ctx_release_buffers(self._h))
        self.table = None
        self.n_kept = self.n_nodes = 0

    def set_tuning(self, l1_chunks: int = 0, l1_bits: int = 0, bucket_factor: float = 0.0,
                   ms_scatter: int = 0, ms_fold: int = 0):
        t = Tuning(l1_chunks, l1_bits, bucket_factor, ms_scatter, ms_fold)
        self._check(self._lib.pk_ctx_set_tuning(self._h, ctypes.byref(t)))

    # -- perform_kmer_counting ------------------------------------------------
    def count(self, fasta, k: int, device_ptr: int | None = None, nbytes: int | None = None) -> TableView:
        """Count k-mers.  `fasta` is bytes / a uint8 numpy array (copied to the
        device), or pass device_ptr + nbytes of a device-resident buffer."""
        t = _Table()
        if device_ptr is not None:
            rc = self._lib.pk_count(self._h, ctypes.c_void_p(device_ptr), nbytes, k, 1, ctypes.byref(t))
        else:
            a = _as_u8(fasta)
            rc = self._lib.pk_count(self._h, a.ctypes.data, a.nbytes, k, 0, ctypes.byref(t))
        self._check(rc)
        self.k = k
        self.table = TableView(int(t.n), int(t.words), int(t.k), t.keys or 0, t.counts or 0)
        return self.table

    def count_file(self, path: str, k: int) -> TableView:
        """Count the k-mers of a FASTA file read straight into HBM through
        pinned staging (pk_count_file)."""
        t = _Table()
        self._check(self._lib.pk_count_file(self._h, os.fsencode(path), k, ctypes.byref(t)))
        self.k = k
        self.table = TableView(int(t.n), int(t.words), int(t.k), t.keys or 0, t.counts or 0)
        return self.table

    def count_noalloc(self, device_ptr: int, nbytes: int, k: int, keys_ptr: int, counts_ptr: int,
                      capacity: int) -> int:
        """pk_count's table of a device-resident FASTA written into the
        caller's device arrays (keys: capacity * words uint64, counts:
        capacity int32) with no device allocation; returns the number of
        distinct k-mers.  Inputs must fit the library's static scratch
        (PK_NOALLOC_SCRATCH_BYTES per GPU)."""
        n = ctypes.c_uint64()
        self._check(self._lib.pk_count_noalloc(self._h, ctypes.c_void_p(device_ptr), nbytes, k,
                                               ctypes.c_void_p(keys_ptr), ctypes.c_void_p(counts_ptr),
                                               capacity, ctypes.byref(n)))
        self.table = None
        return int(n.value)

    def count_format(self, data, k: int, fmt: str = "fasta", device_ptr: int | None = None,
                     nbytes: int | None = None) -> TableView:
        """Count k-mers of multi-line FASTA ("fasta"), four-line FASTQ
        ("fastq") or strict single-line FASTA ("fasta-strict"), '\r\n' line
        ends allowed in the first two: the table of pk_count on the input
        rewritten as strict single-line FASTA (see normalize)."""
        f = FORMATS[fmt]
        t = _Table()
        if device_ptr is not None:
            rc = self._lib.pk_count_format(self._h, ctypes.c_void_p(device_ptr), nbytes, k, f, 1, ctypes.byref(t))
        else:
            a = _as_u8(data)
            rc = self._lib.pk_count_format(self._h, a.ctypes.data, a.nbytes, k, f, 0, ctypes.byref(t))
        self._check(rc)
        self.k = k
        self.table = TableView(int(t.n), int(t.words), int(t.k), t.keys or 0, t.counts or 0)
        return self.table

    def normalize(self, data, fmt: str = "fasta") -> bytes:
        """The input rewritten on the device as strict single-line FASTA."""
        a = _as_u8(data)
        nb = ctypes.c_uint64(0)
        self._check(self._lib.pk_normalize(self._h, a.ctypes.data, a.nbytes, FORMATS[fmt], 0, ctypes.byref(nb)))
        out = np.empty(int(nb.value), dtype=np.uint8)
        if nb.value:
            self._check(self._lib.pk_download_normalized(self._h, out.ctypes.data))
        return out.tobytes()

    def count_rank_share(self, fasta, k: int, nranks: int, rank: int, device_ptr: int | None = None,
                         nbytes: int | None = None) -> TableView:
        """Rank `rank`'s share at `nranks` ranks (owner = hash31(min-frequency
        8-mer) % nranks, kmer_counting.cpp:2950-3050)."""
        t = _Table()
        if device_ptr is not None:
            rc = self._lib.pk_count_rank_share(self._h, ctypes.c_void_p(device_ptr), nbytes, k, nranks, rank,
                                               1, ctypes.byref(t))
        else:
            a = _as_u8(fasta)
            rc = self._lib.pk_count_rank_share(self._h, a.ctypes.data, a.nbytes, k, nranks, rank, 0,
                                               ctypes.byref(t))
        self._check(rc)
        self.k = k
        self.table = TableView(int(t.n), int(t.words), int(t.k), t.keys or 0, t.counts or 0)
        return self.table

    def count_rank_share_by_key(self, fasta, k: int, nranks: int, rank: int) -> TableView:
        """Rank `rank`'s share at `nranks` ranks with the -p counters' owner
        rule: retrieve_proc_id(kmer) (hash31 / uhash31 of the k-mer) % nranks."""
        t = _Table()
        a = _as_u8(fasta)
        self._check(self._lib.pk_count_rank_share_by_key(self._h, a.ctypes.data, a.nbytes, k, nranks, rank, 0,
                                                         ctypes.byref(t)))
        self.k = k
        self.table = TableView(int(t.n), int(t.words), int(t.k), t.keys or 0, t.counts or 0)
        return self.table

    def macro_nodes_rank_share(self, nranks: int, rank: int) -> int:
        """The macro nodes rank `rank` holds at `nranks` ranks (node keys it
        owns by retrieve_proc_id); the context's node table becomes the share."""
        m = _MnTable()
        self._check(self._lib.pk_macro_nodes_rank_share(self._h, nranks, rank, ctypes.byref(m)))
        self.n_nodes = int(m.n)
        self.nodes_ptr = m.nodes or 0
        return self.n_nodes

    def lmer_frequency(self, fasta) -> np.ndarray:
        """4^8 L-mer frequencies (Sliding_window_l, kmer_counting.cpp:391-438)."""
        a = _as_u8(fasta)
        out = np.zeros(1 << 16, dtype=np.uint64)
        self._check(self._lib.pk_lmer_frequency(self._h, a.ctypes.data, a.nbytes, 0, out.ctypes.data))
        return out

    # -- Concurrent_Kmer_Partial_Counter::default_count_and_process ------------
    def builder_count(self, fasta, klen: int, read_len: int = 100, kmer_bits: int = 64,
                      delimiter: bytes = b"\n", skip_symbol: bytes = b">",
                      device_ptr: int | None = None, nbytes: int | None = None) -> int:
        """Builder-API count (concurrent_kmer_counter.hpp:2018-2082); returns the
        number of reads counted, the table is self.table."""
        t = _Table()
        nr = ctypes.c_uint64()
        if device_ptr is not None:
            rc = self._lib.pk_builder_count(self._h, ctypes.c_void_p(device_ptr), nbytes, klen, read_len,
                                            kmer_bits, delimiter, skip_symbol, 1, ctypes.byref(t),
                                            ctypes.byref(nr))
        else:
            a = _as_u8(fasta)
            rc = self._lib.pk_builder_count(self._h, a.ctypes.data, a.nbytes, klen, read_len, kmer_bits,
                                            delimiter, skip_symbol, 0, ctypes.byref(t),
                                            ctypes.byref(nr))
        self._check(rc)
        self.k = klen
        self.table = TableView(int(t.n), int(t.words), int(t.k), t.keys or 0, t.counts or 0)
        return int(nr.value)

    def builder_count_windows(self, fasta, klen: int, read_len: int = 100, mode: str = "v0",
                              delimiter: bytes = b"\n", skip_symbol: bytes = b">") -> tuple[int, int, int]:
        """Window count of the implicit builder counters (omp_implicit_v0 /
        _v1, concurrent_kmer_counter.hpp:910-1022,1363-1578): (records,
        windows, distinct klen-byte windows under strncmp)."""
        a = _as_u8(fasta)
        r, w, d = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        m = {"v0": 0, "v1": 1}[mode]
        self._check(self._lib.pk_builder_count_windows(self._h, a.ctypes.data, a.nbytes, klen, read_len, delimiter,
                                                       skip_symbol, m, 0, ctypes.byref(r), ctypes.byref(w),
                                                       ctypes.byref(d)))
        self.table = None
        return int(r.value), int(w.value), int(d.value)

    def table_host(self):
        """(keys[n, words] uint64, counts[n] int32) of the current table."""
        n, w = self.table.n, self.table.words
        keys = np.empty((n, w), dtype=np.uint64)
        counts = np.empty(n, dtype=np.int32)
        self._check(self._lib.pk_download_table(self._h, keys.ctypes.data, counts.ctypes.data))
        return keys, counts

    # -- construct_kmer_histogram ---------------------------------------------
    def cutoff(self, t: int = 21):
        """(hist[t], min_bucket, n_kept).  Raises PakHipError(PK_EEMPTY) when
        nothing survives (the reference aborts there)."""
        tt = 21 if t == 0 else t
        hist = np.zeros(max(tt, 1), dtype=np.uint64)
        mb = ctypes.c_int()
        nk = ctypes.c_uint64()
        rc = self._lib.pk_cutoff(self._h, t, hist.ctypes.data, ctypes.byref(mb), ctypes.byref(nk))
        self.n_kept = int(nk.value)
        self.min_bucket = int(mb.value)
        self.hist = hist
        self._check(rc)
        return hist, int(mb.value), int(nk.value)

    def cutoff_threshold(self, threshold: int) -> int:
        """Keep count > threshold (Pakman_Subgraph::build_graph,
        pakman_graph.hpp:592-610); returns the kept count."""
        nk = ctypes.c_uint64()
        self._check(self._lib.pk_cutoff_threshold(self._h, threshold, ctypes.byref(nk)))
        self.n_kept = int(nk.value)
        return self.n_kept

    def kept_host(self):
        w = self.table.words
        keys = np.empty((self.n_kept, w), dtype=np.uint64)
        counts = np.empty(self.n_kept, dtype=np.int32)
        self._check(self._lib.pk_download_kept(self._h, keys.ctypes.data, counts.ctypes.data))
        return keys, counts

    # -- construct_macro_nodes --------------------------------------------------
    def macro_nodes(self) -> int:
        m = _MnTable()
        self._check(self._lib.pk_macro_nodes(self._h, ctypes.byref(m)))
        self.n_nodes = int(m.n)
        self.nodes_ptr = m.nodes or 0
        return self.n_nodes

    def mn_wiring(self, coverage: int, pre_order: np.ndarray | None = None) -> np.ndarray:
        """Initial wiring of the last macro nodes (wire.cpp:66-121,
        distribute_kmers.h:550-656); pre_order: uint8 [n_nodes, 4] prefix
        bases in order (255 = unused), None = ascending base."""
        out = np.zeros(self.n_nodes, dtype=MNWIRE_DTYPE)
        po = None
        if pre_order is not None:
            po = np.ascontiguousarray(pre_order, dtype=np.uint8)
            assert po.shape == (self.n_nodes, 4)
        self._check(self._lib.pk_mn_wiring(self._h, coverage, po.ctypes.data if po is not None else None,
                                           out.ctypes.data))
        return out

    def nodes_host(self) -> np.ndarray:
        out = np.empty(self.n_nodes, dtype=MNODE_DTYPE)
        if self.n_nodes:
            self._check(self._lib.pk_download_nodes(self._h, out.ctypes.data))
        return out

    def timing(self) -> dict:
        t = _Timing()
        self._check(self._lib.pk_last_timing(self._h, ctypes.byref(t)))
        out = {}
        for f, _ in _Timing._fields_:
            v = getattr(t, f)
            out[f] = list(v) if isinstance(v, ctypes.Array) else v
        return out


# ---------------------------------------------------------------------------
# Batched / served forms (pk_server_*).

@dataclass
class JobResult:
    ticket: int
    status: int
    min_bucket: int
    n_reads: int
    n_instances: int
    n_unique: int
    n_kept: int
    n_nodes: int
    words: int
    t: int
    t_submit: float
    t_start: float
    t_end: float
    est_bytes: int
    message: str


class Server:
    """n_slots contexts on one device serving count [+ cutoff [+ macro nodes]]
    jobs; results stay resident on the device until release()."""

    def __init__(self, device: int = 0, n_slots: int = 4, hbm_budget_bytes: int = 0):
        self._lib = load_library()
        h = ctypes.c_void_p()
        rc = self._lib.pk_server_create(ctypes.byref(h), device, n_slots, hbm_budget_bytes)
        if rc:
            raise PakHipError(rc, "pk_server_create failed")
        self._h = h
        self._keep = {}   # ticket -> host buffers that must outlive the job
        self._t = {}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pk_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise PakHipError(rc, self._lib.pk_server_last_error(self._h).decode(errors="replace"))

    def now(self) -> float:
        return float(self._lib.pk_server_now(self._h))

    @staticmethod
    def _job(fasta, k, t, want_nodes, device_ptr, nbytes):
        j = _Job()
        keep = None
        if device_ptr is not None:
            j.fasta, j.nbytes, j.fasta_on_device = device_ptr, nbytes, 1
        else:
            keep = _as_u8(fasta)
            j.fasta, j.nbytes, j.fasta_on_device = keep.ctypes.data, keep.nbytes, 0
        j.k, j.t, j.want_nodes = k, t, 1 if want_nodes else 0
        return j, keep

    def submit(self, fasta=None, k: int = 32, t: int = 21, want_nodes: bool = True,
               device_ptr: int | None = None, nbytes: int | None = None) -> int:
        j, keep = self._job(fasta, k, t, want_nodes, device_ptr, nbytes)
        tk = ctypes.c_uint64()
        self._check(self._lib.pk_server_submit(self._h, ctypes.byref(j), ctypes.byref(tk)))
        self._keep[tk.value] = keep
        self._t[tk.value] = 21 if t == 0 else t
        return int(tk.value)

    def submit_batch(self, jobs: list[dict]) -> list[int]:
        """jobs: dicts with the keyword arguments of submit().  Enqueued
        largest first; returns the tickets in the order given."""
        n = len(jobs)
        arr = (_Job * n)()
        keeps = []
        for i, d in enumerate(jobs):
            j, keep = self._job(d.get("fasta"), d.get("k", 32), d.get("t", 21), d.get("want_nodes", True),
                                d.get("device_ptr"), d.get("nbytes"))
            arr[i] = j
            keeps.append(keep)
        tks = (ctypes.c_uint64 * n)()
        self._check(self._lib.pk_server_submit_batch(self._h, arr, n, tks))
        out = [int(x) for x in tks]
        for tk, keep, d in zip(out, keeps, jobs):
            self._keep[tk] = keep
            t = d.get("t", 21)
            self._t[tk] = 21 if t == 0 else t
        return out

    def _result(self, tk, r) -> JobResult:
        return JobResult(tk, r.status, r.min_bucket, r.n_reads, r.n_instances, r.table.n, r.kept.n,
                         r.nodes.n, r.table.words, self._t.get(tk, 0), r.t_submit, r.t_start, r.t_end,
                         r.est_bytes, r.message.decode(errors="replace"))

    def wait(self, ticket: int) -> JobResult:
        r = _JobResult()
        self._check(self._lib.pk_server_wait(self._h, ticket, ctypes.byref(r)))
        self._keep.pop(ticket, None)
        return self._result(ticket, r)

    def poll(self, ticket: int) -> JobResult | None:
        r = _JobResult()
        rc = self._lib.pk_server_poll(self._h, ticket, ctypes.byref(r))
        if rc == PK_EAGAIN:
            return None
        self._check(rc)
        self._keep.pop(ticket, None)
        return self._result(ticket, r)

    def download(self, res: JobResult) -> dict:
        w = max(res.words, 1)
        keys = np.empty((res.n_unique, w), dtype=np.uint64)
        counts = np.empty(res.n_unique, dtype=np.int32)
        kk = np.empty((res.n_kept, w), dtype=np.uint64)
        kc = np.empty(res.n_kept, dtype=np.int32)
        hist = np.zeros(max(res.t, 1), dtype=np.uint64)
        nodes = np.empty(res.n_nodes, dtype=MNODE_DTYPE)
        self._check(self._lib.pk_server_download(self._h, res.ticket, keys.ctypes.data, counts.ctypes.data,
                                                 kk.ctypes.data, kc.ctypes.data, hist.ctypes.data,
                                                 nodes.ctypes.data))
        return {"keys": keys, "counts": counts, "kept_keys": kk, "kept_counts": kc, "hist": hist,
                "nodes": nodes}

    def release(self, ticket: int):
        self._check(self._lib.pk_server_release(self._h, ticket))
        self._t.pop(ticket, None)

    def set_tuning(self, entries: list[dict]):
        """entries: dicts with max_bytes, key_words, l1_chunks, l1_bits,
        bucket_factor, ms_scatter (the autotuner's table, pakman_plus_amd.tune)."""
        arr = (_TuneEntry * max(len(entries), 1))()
        for i, e in enumerate(entries):
            arr[i] = _TuneEntry(int(e["max_bytes"]), int(e.get("key_words", 0)),
                                Tuning(int(e.get("l1_chunks", 0)), int(e.get("l1_bits", 0)),
                                       float(e.get("bucket_factor", 0.0)),
                                       int(e.get("ms_scatter", 0)), int(e.get("ms_fold", 0))))
        self._check(self._lib.pk_server_set_tuning(self._h, arr, len(entries)))


def nodes_as_arrays(nodes: np.ndarray, k: int) -> dict:
    """Node records as the arrays the oracle returns (keys[m, words], pre_cnt, suf_cnt)."""
    w = 1 if k <= 32 else 2
    keys = np.empty((nodes.shape[0], w), dtype=np.uint64)
    keys[:, 0] = nodes["key_lo"]
    if w == 2:
        keys[:, 1] = nodes["key_hi"]
    return {"keys": keys, "pre_cnt": np.ascontiguousarray(nodes["pre_cnt"]),
            "suf_cnt": np.ascontiguousarray(nodes["suf_cnt"])}


# ---------------------------------------------------------------------------
# Function-style mirror of the reference's free functions.

def perform_kmer_counting(ctx: Context, read_data, k: int = 32):
    """kmer_counting.cpp:3351: returns the sorted unique (keys, counts) table
    (the reference's kmer_proc_buf)."""
    ctx.count(read_data, k)
    return ctx.table_host()


def construct_kmer_histogram(ctx: Context, num_buckets: int = 21):
    """kmer_histogram.cpp:72: (hist, min_bucket, kept keys, kept counts)."""
    hist, mb, _ = ctx.cutoff(num_buckets)
    keys, counts = ctx.kept_host()
    return hist, mb, keys, counts


def begin_mnode_construction(ctx: Context, num_buckets: int = 21):
    """mn_construct.cpp:479: cutoff + macro nodes; returns the node records."""
    ctx.cutoff(num_buckets)
    ctx.macro_nodes()
    return ctx.nodes_host()
