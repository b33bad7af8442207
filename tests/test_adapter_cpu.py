"""Host parts of the C++ adapter (include/pakhip_builders.hpp) on the CPU:
the reference prefix order against the oracle's reference-order mode (itself
checked against the reference build's node dumps) and visit_value against
the oracle's restatement of mn_construct.cpp:75-79."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from golden_util import SYNTH, read_fasta

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin", "prefix_order_test")


@pytest.fixture(scope="module")
def prefix_bin():
    from pakman_plus_amd import build
    build.build_prefix_order_test()
    return BIN


def _run(prefix_bin, k, kept_keys, node_keys):
    words = kept_keys.shape[1]
    hdr = f"{k} {words} {kept_keys.shape[0]} {node_keys.shape[0]}\n".encode()
    data = hdr + np.ascontiguousarray(kept_keys).tobytes() + np.ascontiguousarray(node_keys).tobytes()
    out = subprocess.run([prefix_bin], input=data, capture_output=True, check=True).stdout.decode().split("\n")
    m = node_keys.shape[0]
    order = np.array([[int(x) for x in ln.split()] for ln in out[:m]], dtype=np.int32).reshape(m, 4)
    visits = [int(x) for x in out[m:] if x.strip()]
    return order, visits


@pytest.mark.parametrize("name,k", [("mid", 32), ("mid", 64), ("rep", 32), ("tiny", 21), ("tiny", 33)])
def test_reference_prefix_order(oracle_mod, prefix_bin, name, k):
    cov, ks, t = SYNTH[name]
    fa = read_fasta(f"syn_{name}.fa.gz")
    keys, counts = oracle_mod.count(fa, k)
    hist, mb, keep, rc = oracle_mod.cutoff(counts, t)
    assert rc == 0
    want = oracle_mod.macro_nodes(keys[keep], counts[keep], k, ref_order=True)
    order, _ = _run(prefix_bin, k, keys[keep], want["keys"])
    assert np.array_equal(order, want["pre_order"].astype(np.int32))
    # the order differs from ascending base order somewhere (it is not trivial)
    canon = oracle_mod.macro_nodes(keys[keep], counts[keep], k, ref_order=False)
    assert np.array_equal(np.sort(np.where(order == 255, 9, order), axis=1),
                          np.sort(np.where(canon["pre_order"] == 255, 9, canon["pre_order"]), axis=1))


def test_visit_value(oracle_mod, prefix_bin):
    kept = np.zeros((1, 1), dtype=np.uint64)
    _, visits = _run(prefix_bin, 32, kept, np.zeros((0, 1), dtype=np.uint64))
    want = [oracle_mod.visit_value(c, cov) for c in range(0, 300, 7) for cov in (1, 3, 30, 80)]
    assert visits == want
