"""Batched and served forms (pk_server_*): every job's results equal the same
job run alone on a context and the CPU oracle, whatever else runs beside it;
admission by HBM estimate; results stay resident until release."""
from __future__ import annotations

import sys
import threading
import time

import numpy as np
import pytest

from golden_util import read_fasta

pytestmark = pytest.mark.gpu


def _p(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mixed_jobs(gen_mod):
    """Uneven small read sets (sizes, read lengths, k, t, nodes on/off) plus an
    input that fails the format check and one whose cutoff keeps nothing."""
    rng = np.random.default_rng(2024)
    jobs = []
    for i in range(14):
        G = int(np.exp(rng.uniform(np.log(5e3), np.log(6e4))))
        L = int(rng.choice([75, 100, 150, 250]))
        cov = float(rng.uniform(20, 60))
        spec = gen_mod.ReadSetSpec(genome_len=G, read_len=L, coverage=cov, err=0.005,
                                   genome_seed=5000 + i, read_seed=6000 + i)
        fa = gen_mod.make_fasta(spec).tobytes()
        k = int(rng.choice([32, 64, 21, 45]))
        t = int(rng.choice([21, 15, 0]))
        jobs.append({"fasta": fa, "k": k, "t": t, "want_nodes": bool(i % 5 != 3)})
    jobs.append({"fasta": read_fasta("read_test1_raw.fa"), "k": 32, "t": 21, "want_nodes": True})
    jobs.append({"fasta": read_fasta("ref_test_two_reads.fa"), "k": 32, "t": 21, "want_nodes": True})
    return jobs


def _alone(ctx, job):
    """The job on a plain context: (status, table, hist, mb, kept, nodes)."""
    from pakman_plus_amd import api
    try:
        ctx.count(job["fasta"], job["k"])
    except api.PakHipError as e:
        return {"status": e.code}
    keys, counts = ctx.table_host()
    out = {"status": 0, "keys": keys, "counts": counts}
    try:
        hist, mb, nk = ctx.cutoff(job["t"])
    except api.PakHipError as e:
        out.update(status=e.code, hist=ctx.hist, mb=ctx.min_bucket)
        return out
    out.update(hist=hist, mb=mb)
    out["kept_keys"], out["kept_counts"] = ctx.kept_host()
    if job["want_nodes"]:
        ctx.macro_nodes()
        out["nodes"] = ctx.nodes_host()
    return out


def _compare(server, res, want, oracle_mod, job):
    from pakman_plus_amd import api
    assert res.status == want["status"], (res.status, want["status"], res.message)
    if want["status"] in (api.PK_EFORMAT, api.PK_EINVAL):
        return
    got = server.download(res)
    assert np.array_equal(got["keys"], want["keys"]) and np.array_equal(got["counts"], want["counts"])
    assert np.array_equal(got["hist"], want["hist"]) and res.min_bucket == want["mb"]
    ok, oc = oracle_mod.count(job["fasta"], job["k"])
    assert np.array_equal(got["keys"], ok) and np.array_equal(got["counts"], oc)
    if want["status"] == api.PK_EEMPTY:
        assert res.n_kept == 0 and res.n_nodes == 0
        return
    assert np.array_equal(got["kept_keys"], want["kept_keys"])
    assert np.array_equal(got["kept_counts"], want["kept_counts"])
    if job["want_nodes"]:
        assert res.n_nodes == want["nodes"].shape[0]
        assert got["nodes"].tobytes() == want["nodes"].tobytes()
    else:
        assert res.n_nodes == 0


def test_batch_equals_alone_and_oracle(oracle_mod, gen_mod):
    _p("case: server batch of 16 mixed jobs, 4 slots")
    from pakman_plus_amd import api
    jobs = _mixed_jobs(gen_mod)
    with api.Context(0) as ctx:
        alone = [_alone(ctx, j) for j in jobs]
    with api.Server(0, n_slots=4) as srv:
        tickets = srv.submit_batch(jobs)
        results = [srv.wait(tk) for tk in tickets]
        for job, res, want in zip(jobs, results, alone):
            _compare(srv, res, want, oracle_mod, job)
        # largest first
        starts = sorted(results, key=lambda r: r.t_start)
        sizes = [len(jobs[tickets.index(r.ticket)]["fasta"]) for r in starts]
        assert sizes[0] == max(sizes)
        for tk in tickets:
            srv.release(tk)
        with pytest.raises(api.PakHipError) as e:
            srv.wait(tickets[0])
        assert e.value.code == api.PK_EINVAL


def test_batch_with_forced_tuning_equals_alone(oracle_mod, gen_mod):
    """A tuning table that forces the persistent partition scatter and tiny
    buckets (three partition levels, built-in level-1 width) leaves every
    job's results equal to its plain run and the oracle, with four jobs on
    the GPU at once."""
    _p("case: server batch with forced tuning, 4 slots")
    from pakman_plus_amd import api
    jobs = _mixed_jobs(gen_mod)[:8]
    with api.Context(0) as ctx:
        alone = [_alone(ctx, j) for j in jobs]
    table = [{"max_bytes": 1 << 40, "key_words": w, "l1_chunks": 1, "l1_bits": 0,
              "bucket_factor": 0.005, "ms_scatter": 1} for w in (1, 2)]
    with api.Server(0, n_slots=4) as srv:
        srv.set_tuning(table)
        tickets = srv.submit_batch(jobs)
        results = [srv.wait(tk) for tk in tickets]
        for job, res, want in zip(jobs, results, alone):
            _compare(srv, res, want, oracle_mod, job)
        for tk in tickets:
            srv.release(tk)


def test_served_stream_with_device_inputs(oracle_mod, gen_mod):
    """Jobs arrive from another thread while others run (device-resident
    inputs); each result equals the job alone."""
    _p("case: server stream, device inputs")
    import torch
    from pakman_plus_amd import api
    jobs = [j for j in _mixed_jobs(gen_mod)[:10]]
    dev = [torch.from_numpy(np.frombuffer(j["fasta"], dtype=np.uint8).copy()).to("cuda:0") for j in jobs]
    torch.cuda.synchronize()
    with api.Context(0) as ctx:
        alone = [_alone(ctx, j) for j in jobs]
    with api.Server(0, n_slots=3) as srv:
        tickets = [None] * len(jobs)

        def feed():
            for i, (j, d) in enumerate(zip(jobs, dev)):
                tickets[i] = srv.submit(k=j["k"], t=j["t"], want_nodes=j["want_nodes"],
                                        device_ptr=d.data_ptr(), nbytes=d.numel())
                time.sleep(0.002)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        pending = set(range(len(jobs)))
        results = {}
        deadline = time.time() + 60
        while pending and time.time() < deadline:
            for i in list(pending):
                r = srv.poll(tickets[i])
                if r is not None:
                    results[i] = r
                    pending.discard(i)
            time.sleep(0.001)
        assert not pending
        for i, j in enumerate(jobs):
            _compare(srv, results[i], alone[i], oracle_mod, j)
            assert results[i].t_submit <= results[i].t_start <= results[i].t_end
            srv.release(tickets[i])


def test_admission_budget_serialises(gen_mod):
    """A budget smaller than any job's estimate admits one job at a time."""
    _p("case: server admission with a tiny budget")
    from pakman_plus_amd import api
    jobs = _mixed_jobs(gen_mod)[:6]
    with api.Server(0, n_slots=4, hbm_budget_bytes=1) as srv:
        tickets = srv.submit_batch(jobs)
        res = sorted((srv.wait(tk) for tk in tickets), key=lambda r: r.t_start)
        for a, b in zip(res, res[1:]):
            assert b.t_start >= a.t_end
        for r in res:
            assert r.status in (api.PK_OK, api.PK_EEMPTY) and r.est_bytes > 1
            srv.release(r.ticket)


def test_server_tuning_entries_checked():
    """Out-of-range tuning entries are refused as a whole."""
    from pakman_plus_amd import api
    _p("case: server tuning entries")
    with api.Server(0, n_slots=1) as srv:
        good = {"max_bytes": 1 << 30, "key_words": 1, "l1_chunks": 2, "l1_bits": 10, "bucket_factor": 2.2}
        srv.set_tuning([good])
        for bad in ({"l1_bits": 11}, {"l1_chunks": 65}, {"key_words": 3}, {"bucket_factor": float("nan")},
                    {"ms_scatter": 3}):
            with pytest.raises(api.PakHipError) as e:
                srv.set_tuning([good, dict(good, **bad)])
            assert e.value.code == api.PK_EINVAL
        srv.set_tuning([])
