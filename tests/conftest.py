import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

GOLDEN = os.path.join(HERE, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def random_graph(N, E, seed, wmax=1, self_loops=False):
    """Random directed multigraph-free edge list for small parity cases."""
    rng = np.random.default_rng(seed)
    seen = set()
    src, dst = [], []
    while len(src) < E:
        a, b = int(rng.integers(N)), int(rng.integers(N))
        if (a == b and not self_loops) or (a, b) in seen:
            continue
        seen.add((a, b))
        src.append(a)
        dst.append(b)
    w = rng.integers(1, wmax + 1, size=E) if wmax > 1 else None
    from gsap_amd import Graph
    return Graph.from_edges(N, src, dst, w)


def tsv_graph(cat, size="1k"):
    from gsap_amd import read_tsv
    n = {"1k": "1000", "5k": "5000"}[size]
    base = os.path.join(GOLDEN, "graphchallenge", cat, f"{n}_nodes")
    return read_tsv(base + ".tsv", base + "_truePartition.tsv")


def hub_graph(seed=5, N=600, hub_deg=300):
    """Random graph plus two hubs with many in/out neighbours (large LDS tables; beyond 1024
    neighbours, the global hash-table pool)."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    seen = set()

    def add(a, b):
        if a != b and (a, b) not in seen:
            seen.add((a, b)); src.append(a); dst.append(b)
    for _ in range(3000):
        add(int(rng.integers(N)), int(rng.integers(N)))
    for h in (3, 250):
        for t in rng.choice(N, hub_deg, replace=False):
            add(h, int(t)); add(int(t), h)
    from gsap_amd import Graph
    return Graph.from_edges(N, src, dst)


def with_self_loops(g, step=3, wmax=2, seed=0):
    """g plus a self loop (weight 1..wmax) on vertices 0, step, 2 step, ..."""
    from gsap_amd import Graph
    rng = np.random.default_rng(seed)
    src = list(np.repeat(np.arange(g.N), np.diff(g.out_ptr)))
    dst = list(g.out_adj)
    w = list(g.out_w)
    for v in range(0, g.N, step):
        src.append(v); dst.append(v); w.append(int(rng.integers(1, wmax + 1)))
    return Graph.from_edges(g.N, src, dst, np.array(w, np.int32))
