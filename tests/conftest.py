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
