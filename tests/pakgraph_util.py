"""Restatement of the PaK-graph builder's node assembly (Pakman_Subgraph::
build_graph and node_t::rebuild_prefix_suffix_connections,
src/pakman_graph.hpp:80-108,357-408,559-673) from an oracle table, in the
text form tests/cpp/builder_adapter_test.cpp prints.  Parity of this builder
is pinned only by this restatement: the reference's builder does not compile
(incremental_pakgraph.hpp is absent from the reference tree)."""
from __future__ import annotations

import math

BASES = "ACTG"


def _sort_key(e):
    base, count, visit, term = e
    # visit desc, count desc, non-terminal first, shorter string, symbol value
    return (-visit, -count, term, 1 if base else 0, BASES.index(base) if base else 0)


def pakgraph_lines(oracle_mod, fa: bytes, k: int, threshold: int, cov: int) -> list[str]:
    keys, counts = oracle_mod.count(fa, k)
    keep = counts > threshold
    if not keep.any():
        return []
    mn = oracle_mod.macro_nodes(keys[keep], counts[keep], k)
    lines = []
    for i in range(mn["keys"].shape[0]):
        pre = [("", 0, 0, True)] + [(BASES[b], int(c), math.ceil(int(c) / cov), False)
                                    for b, c in enumerate(mn["pre_cnt"][i]) if c]
        suf = [("", 0, 0, True)] + [(BASES[b], int(c), math.ceil(int(c) / cov), False)
                                    for b, c in enumerate(mn["suf_cnt"][i]) if c]
        pre = [list(x) for x in sorted(pre, key=_sort_key)]
        suf = [list(x) for x in sorted(suf, key=_sort_key)]
        tp, ts = sum(p[2] for p in pre), sum(s[2] for s in suf)
        if ts < tp:
            suf[-1][2], suf[-1][1] = tp - ts, 1
        elif ts > tp:
            pre[-1][2], pre[-1][1] = ts - tp, 1
        wires = [[] for _ in pre]
        pi = si = pu = su = 0
        while pi < len(pre) and si < len(suf):
            d = min(pre[pi][2] - pu, suf[si][2] - su)
            pu += d
            su += d
            wires[pi].append((si, d, su))
            if pu == pre[pi][2]:
                pi, pu = pi + 1, 0
            if su == suf[si][2]:
                si, su = si + 1, 0
        lo = int(mn["keys"][i, 0])
        hi = int(mn["keys"][i, 1]) if mn["keys"].shape[1] == 2 else 0
        s = f"{lo} {hi} |"
        for p, w in zip(pre, wires):
            s += f" {p[0] or '-'}:{p[1]}:{p[2]}:{int(p[3])}" + "".join(f":{a}/{b}/{c}" for a, b, c in w)
        s += " |"
        for x in suf:
            s += f" {x[0] or '-'}:{x[1]}:{x[2]}:{int(x[3])}"
        lines.append(s)
    return lines
