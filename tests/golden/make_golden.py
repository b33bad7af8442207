"""Regenerate the golden fixtures under tests/golden/ (run in a container that
has the reference tree at /root/reference and MPICH at /opt/conda).

Fixtures written:
  ref_test_two_reads.fa   the two 100 bp reads of the reference's unit test
                          (src/concurrent_kmer_counter_test.cpp:121-126); the
                          test asserts sum(freq) == reads*(100-k+1) == 138 at k=32
  read_test1.fa           READ_TEST1 (same file :5-44) without its trailing
                          blank line; read_test1_raw.fa keeps it (the main
                          path's parser rejects it, src/kmer_counting.cpp:2928)
  syn_<name>.fa.gz        small seeded read sets (pakman_plus_amd.gen)
  syn_<name>_k<K>.ref.txt.gz
                          output of oracle/_ref/pakref_k<K> (the reference's
                          own kmer_histogram.cpp / mn_construct.cpp / wire.cpp
                          and compaction + walk, OMP_NUM_THREADS=1) on that
                          input: stats lines, the macro nodes after initial
                          wiring (prefixes/suffixes in list order with
                          (count, visit, terminal), wiring), sorted contigs.
"""
from __future__ import annotations

import gzip
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

REF_TEST = "/root/reference/src/concurrent_kmer_counter_test.cpp"

# name -> (generator spec kwargs, coverage flag, list of k, -t)
SYNTH = {
    "tiny": (dict(genome_len=1500, read_len=100, coverage=200.0, err=0.005,
                  genome_seed=9001, read_seed=9002), 200, [21, 32, 33, 64], 21),
    "mid": (dict(genome_len=5000, read_len=150, coverage=120.0, err=0.01,
                 genome_seed=9011, read_seed=9012), 120, [32, 64], 21),
    "rep": (dict(genome_len=6000, read_len=100, coverage=150.0, err=0.005, rep_families=3,
                 rep_min=300, rep_max=600, rep_frac=0.3, rep_div=0.02,
                 genome_seed=9021, read_seed=9022), 150, [32, 64], 15),
}


def c_string_literals(text: str) -> str:
    """Concatenate the C string literals of `text`, resolving \\n escapes."""
    parts = re.findall(r'"((?:[^"\\]|\\.)*)"', text)
    return "".join(p.replace("\\n", "\n") for p in parts)


def extract_ref_test_inputs() -> None:
    src = open(REF_TEST).read()
    m = re.search(r"READ_TEST1\s*=\s*(.*?);", src, re.S)
    read_test1 = c_string_literals(m.group(1))
    m2 = re.search(r"std::make_tuple\(\s*(.*?),\s*32\s*,\s*4\)", src, re.S)
    two = c_string_literals(m2.group(1))
    with open(os.path.join(HERE, "read_test1_raw.fa"), "w") as f:
        f.write(read_test1)
    with open(os.path.join(HERE, "read_test1.fa"), "w") as f:
        f.write(read_test1.rstrip("\n") + "\n")
    with open(os.path.join(HERE, "ref_test_two_reads.fa"), "w") as f:
        f.write(two)


def run_ref(k: int, fasta: str, cov: int, t: int, read_len: int) -> str:
    exe = os.path.join(ROOT, "oracle", "_ref", f"pakref_k{k}")
    work = tempfile.mkdtemp()
    try:
        env = dict(os.environ, OMP_NUM_THREADS="1", PAKREF_DUMP=os.path.join(work, "d"))
        rl = min(max(read_len, 100), 250)
        proc = subprocess.run([exe, "-f", fasta, "-c", str(cov), "-r", str(rl), "-t", str(t),
                               "-n", "1000"], cwd=work, env=env, capture_output=True, text=True,
                              timeout=600)
        out = proc.stdout
        lines = ["# stats", f"exit {0 if proc.returncode == 0 else 'abort'}"]
        for ln in out.splitlines():
            if ln.startswith("Total") or ln.startswith("K-mer size"):
                lines.append(ln.strip())
        if proc.returncode != 0:
            return "\n".join(lines) + "\n"
        lines.append("# wired")
        lines += open(os.path.join(work, "d_wired.txt")).read().splitlines()
        lines.append("# contigs")
        contigs = []
        for fn in ("contigs_out_r0.fa", "contigs_par_r0.fa"):
            p = os.path.join(work, fn)
            if os.path.exists(p):
                contigs += [ln for ln in open(p).read().splitlines() if ln and not ln.startswith(">")]
        lines += sorted(contigs)
        return "\n".join(lines) + "\n"
    finally:
        shutil.rmtree(work)


def main() -> None:
    from pakman_plus_amd import build, gen
    build.build_gen()
    subprocess.check_call(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")])
    extract_ref_test_inputs()
    for name, (spec_kw, cov, ks, t) in SYNTH.items():
        spec = gen.ReadSetSpec(**spec_kw)
        fa = bytes(gen.make_fasta(spec))
        gz = os.path.join(HERE, f"syn_{name}.fa.gz")
        with gzip.GzipFile(gz, "wb", mtime=0) as f:
            f.write(fa)
        with tempfile.NamedTemporaryFile(suffix=".fa", delete=False) as tf:
            tf.write(fa)
            path = tf.name
        try:
            for k in ks:
                txt = run_ref(k, path, cov, t, spec.read_len)
                with gzip.GzipFile(os.path.join(HERE, f"syn_{name}_k{k}.ref.txt.gz"), "wb",
                                   mtime=0) as f:
                    f.write(txt.encode())
                print(name, k, "ok", flush=True)
        finally:
            os.unlink(path)
    # the two-read test and READ_TEST1 through the reference pipeline (k=32)
    for name, cov in (("ref_test_two_reads", 1), ("read_test1", 1)):
        txt = run_ref(32, os.path.join(HERE, f"{name}.fa"), cov, 21, 100)
        with open(os.path.join(HERE, f"{name}_k32.ref.txt"), "w") as f:
            f.write(txt)


if __name__ == "__main__":
    main()
