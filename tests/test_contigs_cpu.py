"""Contig parity of the node records: the library's MN_map filler
(include/pakhip_mnmap.hpp) fills the reference's own MacroNode map from
pk_mnode records with the reference prefix order, and the reference's own
wiring, iterative compaction and walk (src/wire.cpp, src/iterative_compaction.cpp,
src/merge_mn.cpp, src/local_compaction.cpp, src/walk.cpp, compiled from the
reference tree into oracle/_ref by oracle/build_ref.sh) run on it.

The records are the oracle's; the GPU tests show the device's records equal
them bit for bit (test_gpu_parity.py).  Expected: the wired node dump and
the contigs of the reference's own run from the same FASTA (the fixtures'
"# wired" and "# contigs" sections, tests/golden/make_golden.py).

Also checked: planted mistakes (two prefixes swapped, one suffix count
raised) change the wiring dump, so the comparison can fail."""
from __future__ import annotations

import gzip
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

from golden_util import GOLDEN, SYNTH, read_fasta, read_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MNODE = np.dtype([("key_lo", "<u8"), ("key_hi", "<u8"), ("pre_cnt", "<i4", (4,)), ("suf_cnt", "<i4", (4,))])


def _ref_exe(k):
    p = os.path.join(ROOT, "oracle", "_ref", f"pakref_k{k}")
    if not os.path.exists(p):
        pytest.skip(f"{p} not built (needs the reference tree and MPICH: oracle/build_ref.sh)")
    return p


def _records(nodes: dict) -> np.ndarray:
    m = nodes["keys"].shape[0]
    rec = np.zeros(m, dtype=MNODE)
    rec["key_lo"] = nodes["keys"][:, 0]
    if nodes["keys"].shape[1] == 2:
        rec["key_hi"] = nodes["keys"][:, 1]
    rec["pre_cnt"] = nodes["pre_cnt"]
    rec["suf_cnt"] = nodes["suf_cnt"]
    return rec


def _write_records(path, rec: np.ndarray, order: np.ndarray, k: int):
    with open(path, "wb") as f:
        f.write(b"PKMN" + struct.pack("<Qi", rec.shape[0], k))
        f.write(rec.tobytes())
        f.write(np.ascontiguousarray(order, dtype=np.uint8).tobytes())


def _run_filled(k, rec, order, cov, t, work, walk=True):
    """Run the reference's wiring (+ compaction + walk) on the filled MN_map:
    (wired dump lines, sorted contigs)."""
    path = os.path.join(work, "nodes.bin")
    _write_records(path, rec, order, k)
    env = dict(os.environ, OMP_NUM_THREADS="1", PAKREF_DUMP=os.path.join(work, "d"), PAKREF_NODES=path)
    if not walk:
        env["PAKREF_NO_WALK"] = "1"
    proc = subprocess.run([_ref_exe(k), "-f", "/dev/null", "-c", str(cov), "-r", "100", "-t", str(t),
                           "-n", "1000"], cwd=work, env=env, capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0, proc.stderr[-2000:]
    wired = open(os.path.join(work, "d_wired.txt")).read().splitlines()
    contigs = []
    for fn in ("contigs_out_r0.fa", "contigs_par_r0.fa"):
        p = os.path.join(work, fn)
        if os.path.exists(p):
            contigs += [ln for ln in open(p).read().splitlines() if ln and not ln.startswith(">")]
    return wired, sorted(contigs)


def _wired_section(name):
    p = os.path.join(GOLDEN, name)
    with gzip.open(p, "rt") as f:
        lines = f.read().splitlines()
    a = lines.index("# wired")
    b = lines.index("# contigs")
    return lines[a + 1:b]


def _oracle_records(oracle_mod, fa, k, t):
    keys, counts = oracle_mod.count(fa, k)
    _, _, keep, rc = oracle_mod.cutoff(counts, t)
    assert rc == 0
    nodes = oracle_mod.macro_nodes(keys[keep], counts[keep], k, ref_order=True)
    return _records(nodes), nodes["pre_order"]


@pytest.mark.parametrize("name,k,cov,t", [(n, k, c, t) for n, (c, ks, t) in SYNTH.items() for k in ks])
def test_filled_mn_map_gives_reference_contigs(oracle_mod, name, k, cov, t):
    _ref_exe(k)
    fa = read_fasta(f"syn_{name}.fa.gz")
    ref = read_ref(f"syn_{name}_k{k}.ref.txt.gz")
    rec, order = _oracle_records(oracle_mod, fa, k, t)
    with tempfile.TemporaryDirectory() as work:
        wired, contigs = _run_filled(k, rec, order, cov, t, work)
    assert wired == _wired_section(f"syn_{name}_k{k}.ref.txt.gz")
    assert ref.contigs, "the fixture holds contigs"
    assert contigs == sorted(ref.contigs)


def test_planted_record_mistake_changes_the_wiring(oracle_mod):
    name, k = "rep", 32
    cov, _, t = SYNTH[name]
    _ref_exe(k)
    fa = read_fasta(f"syn_{name}.fa.gz")
    rec, order = _oracle_records(oracle_mod, fa, k, t)
    want = _wired_section(f"syn_{name}_k{k}.ref.txt.gz")
    # a node with two prefixes in the reference order: swapping them changes
    # the dump's prefix lists
    two = np.flatnonzero(order[:, 1] != 255)
    assert two.size
    i = int(two[0])
    swapped = order.copy()
    swapped[i, 0], swapped[i, 1] = order[i, 1], order[i, 0]
    with tempfile.TemporaryDirectory() as work:
        wired, _ = _run_filled(k, rec, swapped, cov, t, work, walk=False)
    assert wired != want
    # one suffix count raised by a few visits: the visit values and the wiring
    # change (the walk is skipped: the reference asserts on such a graph)
    bumped = rec.copy()
    j = int(np.flatnonzero(bumped["suf_cnt"][i] > 0)[0])
    bumped["suf_cnt"][i, j] += 7 * cov
    with tempfile.TemporaryDirectory() as work:
        wired, _ = _run_filled(k, bumped, order, cov, t, work, walk=False)
    assert wired != want


def test_mn_map_fill_time(tmp_path):
    """Host cost of materialising the reference's MN_map from node records
    (the filler, pakhip::fill_MN_map, in the reference-built harness): 2e6
    synthetic records with 1-3 prefixes and suffixes each, k = 64.  Printed
    for the design notes; asserted only to finish."""
    k = 64
    exe = _ref_exe(k)
    rng = np.random.default_rng(3)
    m = 2_000_000
    rec = np.zeros(m, dtype=MNODE)
    keys = np.sort(rng.integers(0, 1 << 62, m, dtype=np.int64).astype(np.uint64))
    rec["key_lo"] = keys
    rec["key_hi"] = keys >> np.uint64(3)
    present = rng.random((m, 4)) < 0.45
    present[:, 0] = True
    rec["pre_cnt"] = np.where(present, rng.integers(1, 200, (m, 4)), 0)
    rec["suf_cnt"] = np.where(np.roll(present, 1, axis=1), rng.integers(1, 200, (m, 4)), 0)
    # each node's present prefixes in ascending base order, then 255
    idx = np.argsort(~present, axis=1, kind="stable")
    order = np.where(np.take_along_axis(present, idx, 1), idx, 255).astype(np.uint8)
    path = tmp_path / "nodes.bin"
    _write_records(str(path), rec, order, k)
    env = dict(os.environ, OMP_NUM_THREADS="1", PAKREF_NODES=str(path), PAKREF_FILL_ONLY="1",
               PAKREF_DUMP=str(tmp_path / "d"))
    proc = subprocess.run([exe, "-f", "/dev/null", "-c", "50", "-r", "150", "-t", "21", "-n", "1000"],
                          cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
    assert proc.returncode == 0, proc.stderr[-2000:]
    line = [ln for ln in proc.stdout.splitlines() if ln.startswith("MN_map fill")]
    assert line
    print(line[0], f"({m} records)")
