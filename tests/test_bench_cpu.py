"""Host logic of bench.py and the tuning table (CPU only)."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_algorithmic_bytes_formula():
    sys.path.insert(0, ROOT)
    import bench
    # bacterial workload figures: S, N, U, V, M at k = 32 (W = 8)
    S, N, U, V, M, W = 407368890, 253920000, 77049109, 9199758, 9199796, 8
    ab = bench.alg_bytes(N, U, V, M, S, W)
    want = (S + 2 * N * W + U * (W + 4) + U * 4 + U * (W + 4) + V * (W + 4) + 2 * V * (W + 4)
            + M * (W + 32))
    assert ab["pipeline"] == want
    assert ab["l1_scatter"] == S + N * W and ab["sort"] == N * W + U * (W + 4)
    assert 7.0e9 < want < 7.6e9


def test_tuning_lookup():
    sys.path.insert(0, ROOT)
    from pakman_plus_amd import tune
    entries = [{"max_bytes": 100, "key_words": 1, "l1_chunks": 1, "l1_bits": 8, "bucket_factor": 3.0},
               {"max_bytes": 1000, "key_words": 1, "l1_chunks": 2, "l1_bits": 7, "bucket_factor": 5.0},
               {"max_bytes": 1000, "key_words": 2, "l1_chunks": 4, "l1_bits": 8, "bucket_factor": 1.5}]
    assert tune.for_size(entries, 50, 32)["l1_chunks"] == 1
    assert tune.for_size(entries, 500, 32)["l1_bits"] == 7
    assert tune.for_size(entries, 500, 64)["bucket_factor"] == 1.5
    assert tune.for_size(entries, 5000, 32) == {}
    table = tune.load()
    assert table and {e["key_words"] for e in table} == {1, 2}


def test_bench_fastq_writer_matches_test_writer():
    """bench.fastq_crlf writes the records tests/format_util.to_fastq does."""
    import bench
    from format_util import to_fastq
    from golden_util import read_fasta
    fa = read_fasta("syn_tiny.fa.gz")
    assert bench.fastq_crlf(fa) == to_fastq(fa, b"\r\n")


def test_tuning_merge_keeps_other_key_width():
    """Re-tuning one key width replaces its entries and keeps the other's."""
    from pakman_plus_amd import tune
    old = [{"max_bytes": 100, "key_words": 1, "k": 32, "bucket_factor": 3.0},
           {"max_bytes": 1000, "key_words": 1, "k": 32, "bucket_factor": 5.0},
           {"max_bytes": 100, "key_words": 2, "k": 64, "bucket_factor": 1.5}]
    new = [{"max_bytes": 1000, "key_words": 2, "k": 64, "bucket_factor": 2.2},
           {"max_bytes": 100, "key_words": 2, "k": 64, "bucket_factor": 9.0}]
    ks, merged = tune.merge(new, (64,), old)
    assert ks == [32, 64]
    assert [(e["key_words"], e["max_bytes"], e["bucket_factor"]) for e in merged] == [
        (1, 100, 3.0), (1, 1000, 5.0), (2, 100, 9.0), (2, 1000, 2.2)]
    assert tune.for_size(merged, 50, 64)["bucket_factor"] == 9.0
    assert tune.for_size(merged, 500, 32)["bucket_factor"] == 5.0
