"""Readers for the fixtures in tests/golden/ (see tests/golden/make_golden.py)."""
from __future__ import annotations

import gzip
import os
import re
from dataclasses import dataclass, field

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# name -> (coverage flag, ks, -t); mirrors make_golden.SYNTH
SYNTH = {
    "tiny": (200, [21, 32, 33, 64], 21),
    "mid": (120, [32, 64], 21),
    "rep": (150, [32, 64], 15),
}

BASE_CODE = {"A": 0, "C": 1, "T": 2, "G": 3}


def read_fasta(name: str) -> bytes:
    p = os.path.join(GOLDEN, name)
    if p.endswith(".gz"):
        with gzip.open(p, "rb") as f:
            return f.read()
    with open(p, "rb") as f:
        return f.read()


@dataclass
class RefNode:
    key: tuple[int, int]
    k1mer: str
    prefixes: list  # [(base, count, visit, terminal)]
    suffixes: list
    wiring: list = field(default_factory=list)
    begin: list = field(default_factory=list)


@dataclass
class RefRun:
    aborted: bool
    distinct: int | None
    min_bucket: int | None
    kept: int | None
    macro_nodes: int | None
    nodes: list
    contigs: list


def _ext(tok: list, i: int):
    n = int(tok[i]); i += 1
    out = []
    for _ in range(n):
        out.append((tok[i], int(tok[i + 1]), int(tok[i + 2]), int(tok[i + 3])))
        i += 4
    return out, i


def read_ref(name: str) -> RefRun:
    p = os.path.join(GOLDEN, name)
    opener = gzip.open if p.endswith(".gz") else open
    with opener(p, "rt") as f:
        text = f.read().splitlines()
    run = RefRun(False, None, None, None, None, [], [])
    sec = None
    for ln in text:
        if ln.startswith("# "):
            sec = ln[2:]
            continue
        if sec == "stats":
            if ln.startswith("exit"):
                run.aborted = ln.split()[1] != "0"
            m = re.search(r"distinct k-mer entries across all proc's: (\d+)", ln)
            if m:
                run.distinct = int(m.group(1))
            m = re.search(r"min_bucket\((\d+)\)\) k-mer entries across all proc's: (\d+)", ln)
            if m:
                run.min_bucket, run.kept = int(m.group(1)), int(m.group(2))
            m = re.search(r"number of macro_nodes across all proc's: (\d+)", ln)
            if m:
                run.macro_nodes = int(m.group(1))
        elif sec == "wired":
            tok = ln.split()
            key = (int(tok[1], 16), int(tok[2], 16))
            i = 4
            assert tok[i] == "P"
            pre, i = _ext(tok, i + 1)
            assert tok[i] == "S"
            suf, i = _ext(tok, i + 1)
            assert tok[i] == "W"
            nw = int(tok[i + 1]); i += 2
            wiring = [tuple(int(x) for x in tok[i + 3 * j:i + 3 * j + 3]) for j in range(nw)]
            i += 3 * nw
            assert tok[i] == "B"
            nb = int(tok[i + 1]); i += 2
            begin = [tuple(int(x) for x in tok[i + 2 * j:i + 2 * j + 2]) for j in range(nb)]
            run.nodes.append(RefNode(key, tok[3], pre, suf, wiring, begin))
        elif sec == "contigs":
            run.contigs.append(ln)
    return run


def ref_nodes_as_arrays(run: RefRun, words: int):
    """Macro nodes of a reference run (terminal entries dropped) as arrays in the
    layout of the oracle: keys[m, words], pre_cnt/suf_cnt[m, 4], pre_order[m, 4]."""
    m = len(run.nodes)
    keys = np.zeros((m, words), dtype=np.uint64)
    pre = np.zeros((m, 4), dtype=np.int32)
    suf = np.zeros((m, 4), dtype=np.int32)
    pre_order = np.full((m, 4), 0xFF, dtype=np.uint8)
    suf_order = np.full((m, 4), 0xFF, dtype=np.uint8)
    visits = []
    for i, nd in enumerate(run.nodes):
        keys[i, 0] = nd.key[0]
        if words == 2:
            keys[i, 1] = nd.key[1]
        j = 0
        for b, c, v, term in nd.prefixes:
            if term:
                continue
            pre[i, BASE_CODE[b]] = c
            pre_order[i, j] = BASE_CODE[b]
            visits.append((c, v))
            j += 1
        j = 0
        for b, c, v, term in nd.suffixes:
            if term:
                continue
            suf[i, BASE_CODE[b]] = c
            suf_order[i, j] = BASE_CODE[b]
            visits.append((c, v))
            j += 1
    return {"keys": keys, "pre_cnt": pre, "suf_cnt": suf, "pre_order": pre_order,
            "suf_order": suf_order, "visits": visits}
