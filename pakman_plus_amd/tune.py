"""Autotuner of the counting kernels' launch shapes, per input-size bucket.

For every size bucket (FASTA bytes up to max_bytes) and key width (k = 32:
one word; k = 64: two) it builds a read set of that size (uniform genome,
100 or 150 bp reads, 50x, 0.5% errors), times count + cutoff + macro nodes
on the device for each
candidate (reads per level-1 block, level-1 digit bits, bucket factor = the
inverse of the hash-table load target) and keeps the fastest.  The table is
written as JSON (pakman_plus_amd/tuning.json) and applied by Server.set_tuning
or per context with Context.set_tuning.  Results never depend on the tuning.

  python -m pakman_plus_amd.tune [--out PATH] [--reps 5]
"""
from __future__ import annotations

import argparse
import json
import os
import time

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
# bucket upper bounds (FASTA bytes).  The samples stop near 1 GiB, so larger
# inputs keep the built-in choices, which follow the bucket count (at 5.4 GB,
# k = 64, three partition levels: the built-in 5-bit level 1 beat the 1 GiB
# sample's 10 bits).  A bucket where no candidate beats the built-in choices
# keeps them (zero fields).
BUCKETS = [1 << 21, 1 << 23, 1 << 25, 1 << 27, 1 << 29, 1 << 31]
CANDIDATES = ([{"l1_chunks": c, "l1_bits": 8, "bucket_factor": f}
               for c in (1, 2, 4, 8) for f in (1.5, 2.2, 3.0, 5.0)]
              + [{"l1_chunks": 0, "l1_bits": b, "bucket_factor": 3.0} for b in (6, 7)]
              + [{"l1_chunks": 0, "l1_bits": b, "bucket_factor": f} for b in (9, 10)
                 for f in (1.5, 2.2)])


def _sample(max_bytes: int, prev: int, read_len: int):
    """A read set near the geometric middle of the bucket (prev, max_bytes]."""
    from pakman_plus_amd import gen
    lo = prev if prev else max_bytes // 4
    hi = min(max_bytes, 1 << 30)
    target = int((lo * hi) ** 0.5)
    cov = 50.0
    g = max(1000, int(target / (cov * (1 + 12 / read_len))))
    spec = gen.ReadSetSpec(genome_len=g, read_len=read_len, coverage=cov, err=0.005,
                           genome_seed=9001 + g % 997, read_seed=9002 + g % 991)
    return gen.make_fasta(spec)


def load(path: str = DEFAULT_PATH) -> list[dict]:
    if not os.path.exists(path):
        return []
    with open(path) as f:
        return json.load(f)["entries"]


def for_size(entries: list[dict], nbytes: int, k: int = 32) -> dict:
    """The tuning entry for an input of nbytes at k ({} = built-in choices)."""
    words = 1 if k <= 32 else 2
    for e in sorted(entries, key=lambda e: e["max_bytes"]):
        if nbytes <= e["max_bytes"] and e.get("key_words", 0) in (0, words):
            return {x: e[x] for x in ("l1_chunks", "l1_bits", "bucket_factor", "ms_scatter") if x in e}
    return {}


def tune(path: str = DEFAULT_PATH, reps: int = 5, ks=(32, 64), t: int = 21,
         keep_from: str | None = DEFAULT_PATH) -> dict:
    """Tunes every size bucket for one-word keys (k = 32) and two-word keys
    (k = 64): the hash tables of the two hold different numbers of keys.
    Entries for key widths not tuned now are kept from keep_from."""
    entries = []
    for k in ks:
        entries += _tune_k(reps, k, t)
    ks_all, merged = merge(entries, ks, load(keep_from) if keep_from else [])
    import torch
    out = {"device": torch.cuda.get_device_name(0), "ks": ks_all, "t": t, "reps": reps,
           "entries": merged}
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    return out


def merge(entries: list[dict], ks, previous: list[dict]) -> tuple[list[int], list[dict]]:
    """New entries for the key widths of ks, plus the previous table's
    entries for the other widths: (k values, entries by width and size)."""
    tuned = {1 if k <= 32 else 2 for k in ks}
    kept = [e for e in previous if e.get("key_words") not in tuned]
    ks_all = sorted(set(ks) | {e["k"] for e in kept if "k" in e})
    return ks_all, sorted(kept + entries, key=lambda e: (e.get("key_words", 0), e["max_bytes"]))


def _tune_k(reps: int, k: int, t: int) -> list[dict]:
    import numpy as np
    import torch
    from pakman_plus_amd import api
    ctx = api.Context(0)
    entries = []
    prev = 0
    for mb in BUCKETS:
        fa = _sample(mb, prev, 100 if k <= 32 else 150)
        dev = torch.from_numpy(fa).to("cuda:0")
        torch.cuda.synchronize()
        timings = {}
        for cand in CANDIDATES:
            ctx.set_tuning(**cand)
            ts = []
            for r in range(reps + 1):
                t0 = time.perf_counter()
                ctx.count(None, k, device_ptr=dev.data_ptr(), nbytes=dev.numel())
                try:
                    ctx.cutoff(t)
                    ctx.macro_nodes()
                except api.PakHipError as e:
                    if e.code != api.PK_EEMPTY:
                        raise
                if r:
                    ts.append(time.perf_counter() - t0)
            timings[json.dumps(cand, sort_keys=True)] = float(np.median(ts))
        best = min(timings, key=timings.get)
        e = dict(json.loads(best), max_bytes=mb, key_words=1 if k <= 32 else 2, k=k,
                 sample_bytes=int(fa.nbytes), ms=timings[best] * 1e3, builtin_ms=None)
        ctx.set_tuning()
        ts = []
        for r in range(reps + 1):
            t0 = time.perf_counter()
            ctx.count(None, k, device_ptr=dev.data_ptr(), nbytes=dev.numel())
            try:
                ctx.cutoff(t)
                ctx.macro_nodes()
            except api.PakHipError:
                pass
            if r:
                ts.append(time.perf_counter() - t0)
        e["builtin_ms"] = float(np.median(ts)) * 1e3
        if e["builtin_ms"] < e["ms"]:
            # no candidate beats the built-in choices: zero fields keep them
            e.update(l1_chunks=0, l1_bits=0, bucket_factor=0.0, ms=e["builtin_ms"], builtin=True)
        entries.append(e)
        print(json.dumps(e), flush=True)
        prev = mb
        del dev
    ctx.close()
    return entries


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--ks", default="32,64", help="k values to tune (one per key width)")
    a = ap.parse_args()
    tune(a.out, a.reps, tuple(int(x) for x in a.ks.split(",")))
