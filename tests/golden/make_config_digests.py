#!/usr/bin/env python3
"""Writes the oracle's results on the C3, C4 and C5 workloads as digests for the GPU tests
(tests/test_gpu_configs.py).  Runs the CPU oracle only (no GPU); the graphs come from the
library's host generator.

  c3_digest.json   full partition of the 200k-vertex graph (seed 1, chain 0): final B, S
                   bits, labels sha256, levels, sweeps, the sha256 of the whole trace
                   (every merge, batch, sweep and golden-section record) and the golden-
                   section records themselves
  c4_digests.json  the 256 batched graphs, each partitioned alone (seed 1, chain 0)
  c5_digests.json  the 2M-vertex graph, chain 0 (seed 1)

Usage: python tests/golden/make_config_digests.py [c3] [c4] [c5] [--threads T]
(about 2 min for c3 on 8 threads, c4 a few minutes on 8 processes, c5 longer).
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, os.path.dirname(TESTS))
sys.path.insert(0, TESTS)

import gsap_amd  # noqa: E402
from gsap_amd import workloads  # noqa: E402
import oracle_lib as O  # noqa: E402
from digest_util import result_digest, trace_digest  # noqa: E402


def c3(threads):
    g = workloads.graph("c3")
    t = time.time()
    lab, rep, tr = O.partition(g, gsap_amd.Options(seed=1), trace_cap=1 << 20, threads=threads)
    assert rep.n_trace == len(tr)
    d = result_digest(lab, rep.B, rep.S, rep.levels, rep.sweeps)
    d.update(N=g.N, E=g.E, merge_phases=rep.merge_phases, batches=rep.batches,
             trace=trace_digest(tr), n_trace=len(tr),
             golden=[[r[1], r[2], r[3], f"{r[4]:016x}", r[6]] for r in tr if r[0] == 4],
             oracle_seconds=round(time.time() - t, 1), threads=threads)
    return d


def _c4_one(g):
    gr = workloads.c4_graph(g)
    lab, rep, _ = O.partition(gr, gsap_amd.Options(seed=1))
    d = result_digest(lab, rep.B, rep.S, rep.levels, rep.sweeps)
    d.update(g=g, N=gr.N, E=gr.E)
    return d


def c4(procs):
    t = time.time()
    with mp.Pool(procs) as pool:
        out = pool.map(_c4_one, range(256), chunksize=1)
    return dict(graphs=out, oracle_seconds=round(time.time() - t, 1), processes=procs)


def c5(threads):
    """Chain 0 of the 2M-vertex graph on every host thread (the GPU test checks chain 1
    through the chain entry point's own invariants and a single-chain call)."""
    g = workloads.graph("c5")
    t = time.time()
    lab, rep, _ = O.partition(g, gsap_amd.Options(seed=1, chain=0), threads=threads)
    d = result_digest(lab, rep.B, rep.S, rep.levels, rep.sweeps)
    d.update(chain=0, N=g.N, E=g.E, oracle_seconds=round(time.time() - t, 1), threads=threads)
    return dict(chains=[d])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", nargs="*", default=["c3", "c4", "c5"])
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    for w in a.which:
        t = time.time()
        d = {"c3": c3, "c4": c4, "c5": c5}[w](a.threads)
        path = os.path.join(HERE, f"{w}_digest{'s' if w != 'c3' else ''}.json")
        with open(path, "w") as f:
            json.dump(d, f, indent=1)
        print(f"{w}: {time.time() - t:.1f} s -> {path}", flush=True)


if __name__ == "__main__":
    main()
