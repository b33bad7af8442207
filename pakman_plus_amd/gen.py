"""Seeded synthetic read sets (strict single-line FASTA) for the workloads.

Thin ctypes wrapper over csrc/pk_gen.cpp (built into _lib/libpkgen.so by
pakman_plus_amd.build).  Seeds follow a fixed rule: genome seed =
1000*config + 1, read seed = 1000*config + 2 (per read set s of a batch:
+ 10*s on both).
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(__file__), "_lib", "libpkgen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run pakman_plus_amd.build.build_all() first")
        lib = ctypes.CDLL(path)
        u64, dbl, i32, u32 = ctypes.c_uint64, ctypes.c_double, ctypes.c_int, ctypes.c_uint32
        lib.pkgen_genome.argtypes = [u64, u64, dbl, i32, i32, i32, dbl, dbl, ctypes.c_void_p]
        lib.pkgen_genome.restype = None
        lib.pkgen_num_reads.argtypes = [u64, u32, dbl]
        lib.pkgen_num_reads.restype = u64
        lib.pkgen_fasta_size.argtypes = [u64, u32]
        lib.pkgen_fasta_size.restype = u64
        lib.pkgen_reads.argtypes = [u64, ctypes.c_void_p, u64, u32, u64, dbl, ctypes.c_void_p]
        lib.pkgen_reads.restype = None
        _LIB = lib
    return _LIB


@dataclass
class ReadSetSpec:
    genome_len: int
    read_len: int
    coverage: float
    err: float
    gc: float = 0.5
    rep_families: int = 0
    rep_min: int = 300
    rep_max: int = 5000
    rep_frac: float = 0.0
    rep_div: float = 0.0
    genome_seed: int = 1
    read_seed: int = 2

    def num_reads(self) -> int:
        return int(_lib().pkgen_num_reads(self.genome_len, self.read_len, self.coverage))

    def fasta_size(self) -> int:
        return int(_lib().pkgen_fasta_size(self.num_reads(), self.read_len))


def make_genome(spec: ReadSetSpec) -> np.ndarray:
    g = np.empty(spec.genome_len, dtype=np.uint8)
    _lib().pkgen_genome(spec.genome_seed, spec.genome_len, spec.gc, spec.rep_families,
                        spec.rep_min, spec.rep_max, spec.rep_frac, spec.rep_div,
                        g.ctypes.data)
    return g


def make_fasta(spec: ReadSetSpec, out: np.ndarray | None = None) -> np.ndarray:
    """Return the FASTA bytes as a uint8 array (optionally written into `out`)."""
    lib = _lib()
    genome = make_genome(spec)
    R = spec.num_reads()
    size = int(lib.pkgen_fasta_size(R, spec.read_len))
    if out is None:
        out = np.empty(size, dtype=np.uint8)
    elif out.nbytes < size:
        raise ValueError("output buffer too small")
    if R:
        lib.pkgen_reads(spec.read_seed, genome.ctypes.data, spec.genome_len, spec.read_len, R,
                        spec.err, out.ctypes.data)
    return out[:size]


# Workload presets (sizes as named in the workload descriptions).
def config_small(scale: float = 1.0) -> ReadSetSpec:
    """1 Mbp uniform genome, 100 bp reads, 30x, 0.5% errors; k=32."""
    return ReadSetSpec(genome_len=int(1_000_000 * scale), read_len=100, coverage=30.0,
                       err=0.005, genome_seed=1001, read_seed=1002)


def config_bacterial(scale: float = 1.0) -> ReadSetSpec:
    """4.6 Mbp genome at 50% GC, 100 bp reads, 80x, 1% errors; k=32."""
    return ReadSetSpec(genome_len=int(4_600_000 * scale), read_len=100, coverage=80.0,
                       err=0.01, gc=0.5, genome_seed=2001, read_seed=2002)


def config_large(scale: float = 1.0) -> ReadSetSpec:
    """100 Mbp genome, 5% from 50 repeat families (300-5000 bp, 2% divergence),
    150 bp reads, 50x, 0.5% errors; k=64."""
    return ReadSetSpec(genome_len=int(100_000_000 * scale), read_len=150, coverage=50.0,
                       err=0.005, rep_families=50, rep_min=300, rep_max=5000, rep_frac=0.05,
                       rep_div=0.02, genome_seed=3001, read_seed=3002)


def config_batched(num_sets: int = 256, scale: float = 1.0, config: int = 4) -> list[ReadSetSpec]:
    """Read sets with genome size log-uniform in [50 kbp, 5 Mbp], integer coverage
    uniform in [20, 100], read length from {75, 100, 150, 250}, 0.5% errors."""
    rng = np.random.default_rng(1000 * config + 3)
    specs = []
    for s in range(num_sets):
        g = int(math.exp(rng.uniform(math.log(5e4), math.log(5e6))) * scale)
        cov = int(rng.integers(20, 101))
        L = int(rng.choice([75, 100, 150, 250]))
        g = max(g, L)
        specs.append(ReadSetSpec(genome_len=g, read_len=L, coverage=float(cov), err=0.005,
                                 genome_seed=1000 * config + 1 + 10 * s,
                                 read_seed=1000 * config + 2 + 10 * s))
    return specs
