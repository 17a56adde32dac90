"""GPU runs of the full configurations against the oracle's committed digests
(tests/golden/make_config_digests.py writes them from the CPU oracle):

  C3  the 200k-vertex graph: the whole trace (every merge phase, nodal batch, sweep and
      golden-section step: proposals, labels, block matrix, delta-S sums, S) and the final
      labels, B and S, bit for bit; the untraced call and the partition of the resident
      graph (no second upload) give the same result
  C4  the 256 batched graphs through gsap_partition_batch_i32: every graph's labels, B, S,
      levels and sweeps equal the oracle's; S equals gsap_mdl_i32 of the final labels
  C5  the 2M-vertex graph with chains 0 and 1 through gsap_partition_chains_i32: chain 0's
      S equals the oracle's; chain 1 equals a single-chain call of the same stream (its
      labels, B, S, levels and sweeps); the kept labels are the lower-S chain's, and their
      description length recomputed by gsap_mdl_i32 is the reported S
"""
import json
import os

import numpy as np
import pytest

import gsap_amd
from digest_util import f64_hex, labels_digest, trace_digest
from gsap_amd import Options, workloads

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def digest(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated")
    with open(path) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def ctx():
    c = gsap_amd.Context(0)
    yield c
    c.close()


def test_c3_partition_matches_oracle(ctx):
    d = digest("c3_digest.json")
    g = workloads.graph("c3")
    assert (g.N, g.E) == (d["N"], d["E"])
    print("C3 traced partition", flush=True)
    lab, rep = ctx.partition(g, Options(seed=1, trace=1))
    tr = ctx.trace()
    golden = [[r[1], r[2], r[3], f"{r[4]:016x}", r[6]] for r in tr if r[0] == 4]
    assert golden == d["golden"]
    assert len(tr) == d["n_trace"] and trace_digest(tr) == d["trace"]
    assert (rep.B, f64_hex(rep.S), labels_digest(lab), rep.levels, rep.sweeps) == \
        (d["B"], d["S"], d["labels"], d["levels"], d["sweeps"])
    print("C3 untraced partition", flush=True)
    lab2, rep2 = ctx.partition(g, Options(seed=1))
    assert labels_digest(lab2) == d["labels"] and f64_hex(rep2.S) == d["S"]
    print("C3 partition of the resident graph", flush=True)
    lab3, rep3 = ctx.partition_resident(g.N, Options(seed=1))
    assert labels_digest(lab3) == d["labels"] and f64_hex(rep3.S) == d["S"]


def test_c4_batch_matches_oracle(ctx):
    d = digest("c4_digests.json")["graphs"]
    print("C4 generate 256 graphs", flush=True)
    graphs = workloads.c4_graphs()
    print("C4 batched partition", flush=True)
    outs, reps = ctx.partition_batch(graphs, Options(seed=1))
    for k, (g, lab, rep, want) in enumerate(zip(graphs, outs, reps, d)):
        assert (want["g"], want["N"], want["E"]) == (k, g.N, g.E)
        assert (rep.B, f64_hex(rep.S), labels_digest(lab), rep.levels, rep.sweeps) == \
            (want["B"], want["S"], want["labels"], want["levels"], want["sweeps"]), k
    for k in range(0, 256, 32):
        assert f64_hex(ctx.mdl(graphs[k], outs[k], reps[k].B)) == d[k]["S"], k


def test_c5_chains_match_oracle(ctx):
    d = digest("c5_digests.json")["chains"][0]
    print("C5 generate", flush=True)
    g = workloads.graph("c5")
    assert (g.N, g.E, d["chain"]) == (d["N"], d["E"], 0)
    print("C5 two chains", flush=True)
    lab, rep, S, best = ctx.partition_chains(g, 2, Options(seed=1))
    assert f64_hex(S[0]) == d["S"]
    assert best == min(range(2), key=lambda k: (S[k], k))
    assert f64_hex(rep.S) == f64_hex(S[best])
    print("C5 chain 1 alone", flush=True)
    lab1, rep1 = ctx.partition(g, Options(seed=1, chain=1))
    assert f64_hex(rep1.S) == f64_hex(S[1])
    if best == 0:
        assert labels_digest(lab) == d["labels"] and rep.B == d["B"]
        assert (rep.levels, rep.sweeps) == (d["levels"], d["sweeps"])
    else:
        assert np.array_equal(lab, lab1) and rep.B == rep1.B
        assert (rep.levels, rep.sweeps) == (rep1.levels, rep1.sweeps)
    print("C5 description length of the kept labels", flush=True)
    assert f64_hex(ctx.mdl(g, lab, rep.B)) == f64_hex(rep.S)
