#!/usr/bin/env python3
"""Throughput of the k-mer counting -> histogram cutoff -> macro-node path on
one MI355X.

One step = pk_count + pk_cutoff + pk_macro_nodes over the whole read set of
the workload, with the FASTA bytes already resident in HBM.  Prints one JSON
line (metric, value in k-mer instances per second, roofline of the dominant
kernel, CPU baseline: the oracle's restatement of the reference at P ranks
on the host's cores, and at one rank, on bounded samples of the same reads).

  python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--workload large]
                  [--extra-workloads bacterial,small]
                  [--forms plain,batched,served,formats] [--sets 256] [--slots 4]

The line's own figures are the large workload's (100 Mbp genome with repeat
families, 150 bp reads, 50x, k = 64); the bacterial and small workloads are
timed the same way and reported under "workloads".  The batched and served
forms run after the plain timed steps, on the 256-set workload (see
bench_batched_served), and are reported under "batched" and "served";
"formats" times the bacterial reads as FASTQ with '\r\n' line ends
(bench_formats).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X
# shares of the reads the CPU baseline takes per workload, (all ranks, one
# rank): each leg about 10-20 s of host time
CPU_FRAC = {"large": (0.1, 0.01), "bacterial": (1.0, 0.1), "small": (1.0, 1.0)}
METRIC = "k-mers/s + % HBM peak, batched/served vs plain; p99 job latency at load"

WORKLOADS = {
    "bacterial": ("config_bacterial", 32,
                  "bacterial: 4.6 Mbp genome at 50% GC, 100 bp reads, 80x, 1% errors, k=32"),
    "small": ("config_small", 32, "small: 1 Mbp genome, 100 bp reads, 30x, 0.5% errors, k=32"),
    "large": ("config_large", 64,
              "large with repeats: 100 Mbp genome (5% from 50 repeat families, 2% divergence), "
              "150 bp reads, 50x, 0.5% errors, k=64"),
}


# single-launch stages timed by HIP events around their kernel -> (kernel
# name, timing field); the further partition levels are added per level from
# the per-level fields (ms_level_*), their scatter kernel by form
KERNELS = {"hist": ("k_l1_hist<{key}, 0, false>", "hist_ms"),
           "l1_scatter": ("k_l1_scatter<{key}, 0, false>", "l1_scatter_ms"),
           "sort": ("k_bucket_hash<{key}>", "sort_ms"),
           "compact": ("k_compact<{key}>", "compact_ms")}
MS_SCATTER = {1: "k_ms_scatter_p", 2: "k_ms_scatter", 3: "k_ms_scatter_seg"}


def alg_bytes(N, U, V, M, S, W):
    """Algorithmic HBM bytes per kernel launch (each datum read or written once)."""
    return {
        "hist": S,                              # read the FASTA
        "l1_scatter": S + N * W,                # read the FASTA, write the instances
        "ms_hist": N * W,                       # read the instances
        "ms_scatter": 2 * N * W,                # read and write the instances
        "sort": N * W + U * (W + 4),            # read the instances, write (key, count)
        "compact": 2 * U * (W + 4),             # read and write the unique table
        "cutoff": U * 4 + U * (W + 4) + V * (W + 4),
        "mn": 2 * V * (W + 4) + M * (W + 32),
        # whole path, one read of the input and one write + read of the instance stream
        "pipeline": (S + 2 * N * W + U * (W + 4) + U * 4 + U * (W + 4) + V * (W + 4)
                     + 2 * V * (W + 4) + M * (W + 32)),
    }


def kernel_times(tm, W):
    """[(stage, kernel, ms per step)] of every single-launch kernel the stage
    timings cover: the four fixed stages, then each further partition level's
    histogram (a k_ms_hist launch unless the level's counts were folded into
    level 1's histogram, form 3) and scatter."""
    key = "unsigned long" if W == 8 else "unsigned __int128"
    out = [(stage, name.format(key=key), float(tm.get(field, 0.0))) for stage, (name, field) in KERNELS.items()]
    levels = min(int(tm.get("ms_levels", 0)), len(tm.get("ms_level_form", [])))
    for lv in range(levels):
        form = int(tm["ms_level_form"][lv])
        if form != 3:
            out.append(("ms_hist", f"k_ms_hist<{key}>", float(tm["ms_level_hist_ms"][lv])))
        out.append(("ms_scatter", f"{MS_SCATTER[form]}<{key}>", float(tm["ms_level_scatter_ms"][lv])))
    return [x for x in out if x[2] > 0]


def dominant_kernel(tm, N, U, V, M, S, W, workload):
    """The kernel with the largest time per step (each is one launch), and its
    achieved algorithmic bandwidth (alg_bytes per launch / time per launch)."""
    ab = alg_bytes(N, U, V, M, S, W)
    stage, kernel, per_ms = max(kernel_times(tm, W), key=lambda x: x[2])
    achieved = ab[stage] / (per_ms * 1e-3)
    return {"bound": "hbm", "kernel": kernel,
            "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": load_traffic(kernel, workload),
            "alg_bytes_per_launch": ab[stage], "avg_ms": per_ms, "launches_per_step": 1,
            "share_of_step": None}


def _read_sample(arr, frac):
    """The first `frac` of the reads of a strict single-line FASTA: (bytes, reads, all reads)."""
    import numpy as np
    nl = np.flatnonzero(arr == 10)
    R = nl.size // 2
    take = max(1, min(R, int(round(R * frac))))
    return arr[:int(nl[2 * take - 1]) + 1], take, R


def cpu_baseline(fa, k, t, frac_ranks, frac_one):
    """The reference's algorithm on the host (the oracle's restatement of
    count -> cutoff -> macro nodes at P ranks, one thread per rank, as
    mpiexec -np P with OMP_NUM_THREADS=1): P = the host cores given to one GPU
    (at most 16) on the first frac_ranks of the reads, and one rank on the
    first frac_one of them."""
    import numpy as np
    from oracle import oracle as O
    O.build()
    arr = np.asarray(fa)
    ranks = max(1, min(16, os.cpu_count() or 1))
    out = {}
    for key, P, frac in (("ranks", ranks, frac_ranks), ("np1", 1, frac_one)):
        sample, take, R = _read_sample(arr, frac)
        t0 = time.perf_counter()
        r = O.pipeline_ranks(sample, k, t, P)
        dt = time.perf_counter() - t0
        out[key] = {"value": r["instances"] / dt, "unit": "k-mers/s", "cores": P, "seconds": dt,
                    "sample": f"first {take} of {R} reads ({frac:.0%}) of the same workload: "
                              f"{r['instances']} k-mer instances, {r['distinct']} distinct, {r['kept']} kept, "
                              f"{r['nodes']} macro nodes; count + cutoff + macro nodes at {P} rank(s)"}
    res = dict(out["ranks"])
    res["kind"] = "port"
    res["np1"] = out["np1"]
    return res


def make_batched_sets(num_sets, threads=8):
    """FASTA bytes of the batched workload's read sets (generated in parallel)."""
    from concurrent.futures import ThreadPoolExecutor
    from pakman_plus_amd import gen
    specs = gen.config_batched(num_sets)
    with ThreadPoolExecutor(threads) as ex:
        fas = list(ex.map(gen.make_fasta, specs))
    return specs, fas


def run_plain_job(ctx, dev, k, t, table=None):
    """One job on a plain context, synchronous: (seconds, instances, U, V, M).
    With a tuning table the job's size entry is applied first (as the
    server does per job)."""
    from pakman_plus_amd import api, tune
    ctx.set_tuning(**tune.for_size(table or [], int(dev.numel()), k))
    t0 = time.perf_counter()
    ctx.count(None, k, device_ptr=dev.data_ptr(), nbytes=dev.numel())
    U = ctx.table.n
    V = M = 0
    try:
        ctx.cutoff(t)
        V = ctx.n_kept
        M = ctx.macro_nodes()
    except api.PakHipError as e:
        if e.code != api.PK_EEMPTY:
            raise
    dt = time.perf_counter() - t0
    return dt, int(ctx.timing()["n_instances"]), U, V, M


def bench_batched_served(args):
    """Batched (256 uneven read sets, k=32, one submission) and served (Poisson
    arrivals of the same sets with k from {32, 64}) against one-at-a-time plain
    runs of the same jobs on one context.  Inputs are resident in HBM."""
    import numpy as np
    import torch
    from pakman_plus_amd import api
    t_gen = time.perf_counter()
    specs, fas = make_batched_sets(args.sets)
    t_gen = time.perf_counter() - t_gen
    devs = [torch.from_numpy(f).to("cuda:0") for f in fas]
    torch.cuda.synchronize()
    nbytes = sum(int(f.nbytes) for f in fas)
    del fas
    W = 8
    out = {}
    # -- plain, one job at a time (k = 32) -------------------------------------
    # warm-up grows the buffers to the largest job (u128 keys at k = 64)
    big = sorted(range(len(devs)), key=lambda i: -devs[i].numel())
    from pakman_plus_amd import tune
    table = tune.load()
    ctx = api.Context(0)
    run_plain_job(ctx, devs[big[0]], 64, args.t)
    plain32 = [run_plain_job(ctx, d, 32, args.t, table) for d in devs]
    plain_s = sum(p[0] for p in plain32)
    N = sum(p[1] for p in plain32)
    alg = sum(alg_bytes(p[1], p[2], p[3], p[4], int(d.numel()), W)["pipeline"] for p, d in zip(plain32, devs))
    # -- batched ------------------------------------------------------------------
    srv = api.Server(0, n_slots=args.slots)
    srv.set_tuning(table)
    jobs = [{"k": 32, "t": args.t, "want_nodes": True, "device_ptr": d.data_ptr(), "nbytes": d.numel()}
            for d in devs]
    warm = srv.submit_batch([dict(jobs[i], k=64) for i in big[: args.slots]])
    for tk in warm:
        srv.release(srv.wait(tk).ticket)
    t0 = time.perf_counter()
    tickets = srv.submit_batch(jobs)
    res = [srv.wait(tk) for tk in tickets]
    batched_s = time.perf_counter() - t0
    same = all(r.n_unique == p[2] and r.n_kept == p[3] and r.n_nodes == p[4] and r.n_instances == p[1]
               for r, p in zip(res, plain32))
    for tk in tickets:
        srv.release(tk)
    out["batched"] = {
        "workload": f"{args.sets} read sets: genome log-uniform 50 kbp-5 Mbp, coverage 20-100x, "
                    f"reads 75/100/150/250 bp, 0.5% errors, k=32, -t {args.t}",
        "sets": args.sets, "fasta_bytes": nbytes, "kmer_instances": N, "slots": args.slots,
        "value": N / batched_s, "unit": "k-mers/s", "seconds": batched_s,
        "plain_value": N / plain_s, "plain_seconds": plain_s, "speedup_vs_plain": plain_s / batched_s,
        "hbm_frac": alg / batched_s / HBM_PEAK, "plain_hbm_frac": alg / plain_s / HBM_PEAK,
        "results_match_plain": same, "generation_s": t_gen,
        "tuning": "autotuned per size (pakman_plus_amd/tuning.json), same table for plain and batched"
                  if table else "built-in",
    }
    # -- served: Poisson arrivals, k from {32, 64}, at several offered loads -------
    rng = np.random.default_rng(5003)
    order = rng.permutation(len(devs))
    ks = rng.choice([32, 64], size=len(devs))
    plain64 = {int(i): run_plain_job(ctx, devs[i], 64, args.t, table) for i in order if ks[i] == 64}
    service = np.array([plain64[int(i)][0] if ks[i] == 64 else plain32[int(i)][0] for i in order])
    inst = np.array([plain64[int(i)][1] if ks[i] == 64 else plain32[int(i)][1] for i in order])
    capacity = len(order) / service.sum()  # plain one-at-a-time jobs per second
    points = []
    for load in args.loads:
        lam = load * capacity
        arrivals = np.cumsum(rng.exponential(1.0 / lam, size=len(order)))
        # plain one-at-a-time FIFO with the measured service times
        fin = np.empty(len(order))
        tcur = 0.0
        for i in range(len(order)):
            tcur = max(tcur, arrivals[i]) + service[i]
            fin[i] = tcur
        plain_lat = fin - arrivals
        plain_span = fin[-1] - arrivals[0]
        # a consumer takes each result as it completes and releases it (a
        # serving loop keeps results only until they are read)
        import queue
        import threading
        tickets: "queue.Queue" = queue.Queue()
        rs = []

        def consume():
            for _ in range(len(order)):
                x = tickets.get()
                rs.append(srv.wait(x))
                srv.release(x)

        th = threading.Thread(target=consume)
        th.start()
        wall0 = time.perf_counter() + 0.01
        for i, j in enumerate(order):
            while time.perf_counter() - wall0 < arrivals[i]:
                pass
            d = devs[int(j)]
            tickets.put(srv.submit(k=int(ks[j]), t=args.t, want_nodes=True, device_ptr=d.data_ptr(),
                                   nbytes=d.numel()))
        th.join()
        lat = np.array([r.t_end - r.t_submit for r in rs])
        span = max(r.t_end for r in rs) - min(r.t_submit for r in rs)
        points.append({
            "load": load, "arrival_rate_per_s": lam,
            "value": float(inst.sum() / span), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "plain_value": float(inst.sum() / plain_span),
            "plain_p50_ms": float(np.percentile(plain_lat, 50) * 1e3),
            "plain_p99_ms": float(np.percentile(plain_lat, 99) * 1e3),
        })
    ctx.close()
    srv.close()
    top = points[-1]
    out["served"] = {
        "workload": "Poisson arrivals of the batched read sets, k drawn from {32, 64}",
        "requests": len(order), "slots": args.slots, "kmer_instances": int(inst.sum()),
        "load_rule": "offered load = arrival rate / the plain one-at-a-time capacity (jobs / sum of "
                     "measured service times); plain = one job at a time on one context, FIFO",
        "plain_capacity_per_s": capacity,
        "unit": "k-mers/s",
        # headline: the highest offered load (past plain capacity)
        "load": top["load"], "value": top["value"], "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"],
        "plain_value": top["plain_value"], "plain_p50_ms": top["plain_p50_ms"], "plain_p99_ms": top["plain_p99_ms"],
        "speedup_vs_plain": top["value"] / top["plain_value"],
        "sweep": points,
    }
    del devs
    torch.cuda.empty_cache()
    return out


def fastq_crlf(fa_bytes: bytes) -> bytes:
    """The reads of a strict single-line FASTA as four-line FASTQ records
    with '\r\n' line ends (constant quality line)."""
    lines = fa_bytes.split(b"\n")
    out = []
    for i in range(0, len(lines) - 1, 2):
        seq = lines[i + 1]
        out.append(b"@" + lines[i][1:] + b"\r\n" + seq + b"\r\n+\r\n" + b"I" * len(seq) + b"\r\n")
    return b"".join(out)


def bench_formats(fa, k, t, steps, warmup, tun, plain_ms):
    """pk_count_format on the same reads written as FASTQ with '\r\n' line
    ends, resident in HBM: the device rewrite alone (pk_normalize) and whole
    steps (pk_count_format + pk_cutoff + pk_macro_nodes)."""
    import numpy as np
    import torch
    from pakman_plus_amd import api
    fq = np.frombuffer(fastq_crlf(fa.tobytes()), dtype=np.uint8)
    dev = torch.from_numpy(fq).to("cuda:0")
    torch.cuda.synchronize()
    ctx = api.Context(0)
    if tun:
        ctx.set_tuning(**tun)
    import ctypes
    nb = ctypes.c_uint64(0)

    def normalize():
        ctx._check(ctx._lib.pk_normalize(ctx._h, ctypes.c_void_p(dev.data_ptr()), int(fq.nbytes),
                                         api.FORMATS["fastq"], 1, ctypes.byref(nb)))
        return ctx.timing()["normalize_ms"]

    def step():
        ctx.count_format(None, k, "fastq", device_ptr=dev.data_ptr(), nbytes=int(fq.nbytes))
        ctx.cutoff(t)
        ctx.macro_nodes()
        return ctx.timing()

    for _ in range(warmup):
        normalize()
        step()
    norm = sorted(normalize() for _ in range(steps))[steps // 2]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tm = step()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    n_in, n_out = int(fq.nbytes), int(nb.value)
    ctx.close()
    del dev
    return {"workload": "the bacterial reads as four-line FASTQ records with \\r\\n line ends",
            "input_bytes": n_in, "rewritten_bytes": n_out,
            "normalize_ms": norm,
            "normalize_alg_bytes": n_in + n_out,
            "normalize_frac_of_peak": (n_in + n_out) / (norm * 1e-3) / HBM_PEAK,
            "step_ms": el * 1e3, "value": int(tm["n_instances"]) / el, "unit": "k-mers/s",
            "plain_step_ms": plain_ms}


def bench_ingest(fa, k, t, steps, tun, plain_ms):
    """Whole steps from a FASTA file (pk_count_file: file -> pinned staging ->
    HBM, the file in the page cache) against the same steps with the FASTA
    already resident in HBM."""
    import tempfile
    from pakman_plus_amd import api
    with tempfile.NamedTemporaryFile(suffix=".fa", delete=False) as f:
        f.write(memoryview(fa))
        path = f.name
    try:
        ctx = api.Context(0)
        if tun:
            ctx.set_tuning(**tun)
        def step():
            ctx.count_file(path, k)
            ctx.cutoff(t)
            ctx.macro_nodes()
            ctx.timing()
        step()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        el = (time.perf_counter() - t0) / steps
        ctx.close()
    finally:
        os.unlink(path)
    extra = max(el * 1e3 - plain_ms, 1e-6)
    return {"workload": "the bacterial reads from a FASTA file in the page cache", "bytes": int(fa.nbytes),
            "step_ms": el * 1e3, "plain_step_ms": plain_ms,
            "file_to_hbm_GBps": int(fa.nbytes) / (extra * 1e-3) / 1e9,
            "note": "file- and PCIe-inclusive; not the line's value"}


def load_traffic(kernel, workload):
    """HBM bytes per launch of the kernel on the workload, from the rocprofv3
    PMC runs (FETCH_SIZE and WRITE_SIZE, each scaled by its reading for a
    plain copy of known size) summarised in profiles/traffic.json; None when
    not measured."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        e = json.load(f).get(f"{kernel} @ {workload}")
    return float(e["bytes"]) if e else None


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device copy of known size, for scaling PMC byte counters."""
    import torch
    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0")
    b = torch.empty_like(a)
    a.fill_(1)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": nbytes, "bytes_written": nbytes, "reps": reps}))


def run_workload(name, steps, warmup, t):
    """Plain steps of one workload, FASTA resident in HBM: its figures for the
    JSON line (and the FASTA bytes, for the CPU baseline and the formats form)."""
    import torch
    from pakman_plus_amd import api, gen, tune
    fn, k, desc = WORKLOADS[name]
    spec = getattr(gen, fn)()
    fa = gen.make_fasta(spec)
    nbytes = int(fa.nbytes)
    dev = torch.from_numpy(fa).to("cuda:0")
    torch.cuda.synchronize()
    ctx = api.Context(0)
    tun = tune.for_size(tune.load(), nbytes, k)
    if tun:
        ctx.set_tuning(**tun)

    def step():
        ctx.count(None, k, device_ptr=dev.data_ptr(), nbytes=nbytes)
        ctx.cutoff(t)
        ctx.macro_nodes()
        return ctx.timing()

    for _ in range(warmup):
        step()
    keys = ("parse_ms", "hist_ms", "scatter_ms", "l1_scatter_ms", "ms_hist_ms", "ms_scatter_ms", "sort_ms",
            "split_ms", "compact_ms", "cutoff_ms", "mn_ms")
    lkeys = ("ms_level_hist_ms", "ms_level_scatter_ms")
    acc = {x: 0.0 for x in keys}
    lacc = {x: [0.0] * 4 for x in lkeys}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tm = step()
        for x in keys:
            acc[x] += tm[x]
        for x in lkeys:
            lacc[x] = [a + b for a, b in zip(lacc[x], tm[x])]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    N = int(tm["n_instances"])
    U, V, M = ctx.table.n, ctx.n_kept, ctx.n_nodes
    W = 8 if k <= 32 else 16
    ms_per_step = elapsed * 1e3 / steps
    avg = {x: acc[x] / steps for x in keys}
    avg.update({"ms_levels": int(tm["ms_levels"]), "ms_scatter_form": int(tm["ms_scatter_form"]),
                "ms_level_form": list(tm["ms_level_form"])})
    avg.update({x: [v / steps for v in lacc[x]] for x in lkeys})
    ab = alg_bytes(N, U, V, M, nbytes, W)
    roof = dominant_kernel(avg, N, U, V, M, nbytes, W, name)
    roof["share_of_step"] = roof["avg_ms"] * roof["launches_per_step"] / ms_per_step
    ctx.close()
    del dev
    torch.cuda.empty_cache()
    out = {
        "value": N * steps / elapsed,
        "ms_per_step": ms_per_step,
        "steps": steps, "warmup": warmup,
        "dtype": "u64" if k <= 32 else "u128 (two u64 words)",
        "config": {"workload": desc, "k": k, "t": t, "reads": int(tm["n_reads"]),
                   "kmer_instances": N, "unique_kmers": U, "kept_kmers": V, "macro_nodes": M,
                   "fasta_bytes": nbytes, "bucket_bits": int(tm["bucket_bits"]),
                   "partition_levels": 1 + int(tm["ms_levels"]),
                   "split_buckets": int(tm["n_split"]), "parallelism": "1 GPU",
                   "tuning": tun or "built-in"},
        "roofline": roof,
        "pipeline": {"alg_bytes": ab["pipeline"], "alg_bytes_per_instance": ab["pipeline"] / max(N, 1),
                     "frac_of_peak": ab["pipeline"] / (ms_per_step * 1e-3) / HBM_PEAK,
                     "stage_ms": {x.replace("_ms", ""): v for x, v in avg.items()},
                     # every single-launch kernel timed per step, longest first
                     "kernels": [{"kernel": kn, "ms": ms, "alg_bytes": ab[sg],
                                  "frac": ab[sg] / (ms * 1e-3) / HBM_PEAK,
                                  "traffic": load_traffic(kn, name)}
                                 for sg, kn, ms in sorted(kernel_times(avg, W), key=lambda x: -x[2])]},
    }
    return out, fa, k, tun


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="large", choices=sorted(WORKLOADS))
    ap.add_argument("--extra-workloads", default="bacterial,small",
                    help="workloads timed after the main one, reported under 'workloads' ('' for none)")
    ap.add_argument("--t", type=int, default=21)
    ap.add_argument("--cpu-frac", type=float, default=None,
                    help="share of the reads the CPU baseline's all-rank leg counts (the one-rank "
                         "leg a tenth of it; default by workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--forms", default="plain,batched,served,formats,ingest",
                    help="plain, plus the batched/served forms on the 256-set workload and the "
                         "FASTQ form of the bacterial reads")
    ap.add_argument("--sets", type=int, default=256)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--loads", type=float, nargs="+", default=[0.5, 0.8, 1.2, 2.0],
                    help="offered loads of the served form, as multiples of the plain capacity")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "exchange of this path is a design, not a build", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return

    main_out, fa, k, tun = run_workload(args.workload, args.steps, args.warmup, args.t)
    line = {
        "metric": METRIC,
        "value": main_out["value"],
        "unit": "k-mers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_out["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": main_out["dtype"],
        "data": "synthetic (seeded generator, pakman_plus_amd.gen)",
        "config": main_out["config"],
        "roofline": main_out["roofline"],
        "pipeline": main_out["pipeline"],
    }
    forms = set(args.forms.split(","))
    extras = [w for w in args.extra_workloads.split(",") if w and w != args.workload]
    main_fa = fa
    if extras:
        line["workloads"] = {}
        for w in extras:
            sub, wfa, wk, wtun = run_workload(w, min(args.steps, 10), min(args.warmup, 2), args.t)
            line["workloads"][w] = sub
            if w == "bacterial" and "formats" in forms:
                line["formats"] = bench_formats(wfa, wk, args.t, min(args.steps, 10), min(args.warmup, 2), wtun,
                                                sub["ms_per_step"])
            if w == "bacterial" and "ingest" in forms:
                line["ingest"] = bench_ingest(wfa, wk, args.t, min(args.steps, 10), wtun, sub["ms_per_step"])
            del wfa
    elif args.workload == "bacterial" and "formats" in forms:
        line["formats"] = bench_formats(fa, k, args.t, args.steps, args.warmup, tun, main_out["ms_per_step"])
    if forms & {"batched", "served"}:
        line.update(bench_batched_served(args))
    if not args.no_cpu_baseline:
        fr, f1 = CPU_FRAC.get(args.workload, (0.25, 0.05))
        if args.cpu_frac is not None:
            fr, f1 = args.cpu_frac, args.cpu_frac / 10
        line["cpu_baseline"] = cpu_baseline(main_fa, k, args.t, fr, f1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
