"""Quality report (gsap_quality_i32, GSAP/evaluate/evaluate.hpp:19-265): pinned against the
reference's own evaluation code compiled from the checkout (oracle/_ref/ref_eval, when it
has been built), against brute-force matching on tiny tables, and against identities."""
import itertools
import math
import os
import subprocess

import numpy as np
import pytest

import gsap_amd
from conftest import ROOT

REF_EVAL = os.path.join(ROOT, "oracle", "_ref", "ref_eval")


def labels(rng, n, B):
    x = rng.integers(0, B, n)
    x[:B] = np.arange(B)  # every label present (the reference indexes labels densely)
    return x.astype(np.int32)


def ref_eval(tmp_path, t, g):
    p = tmp_path / "lab.txt"
    p.write_text(f"{len(t)}\n" + " ".join(map(str, t)) + "\n" + " ".join(map(str, g)) + "\n")
    out = subprocess.run([REF_EVAL, str(p)], capture_output=True, text=True, check=True).stdout
    vals = {}
    for line in out.splitlines():
        k, v = line.rsplit(":", 1)
        vals[k.strip()] = float(v)
    return vals


@pytest.mark.skipif(not os.path.exists(REF_EVAL), reason="reference checker not built here")
@pytest.mark.parametrize("seed,n,Bt,Bg,noise", [(1, 300, 5, 9, 0.3), (2, 2000, 12, 7, 0.5),
                                                 (3, 1500, 20, 20, 0.1), (4, 800, 3, 40, 0.8),
                                                 (5, 3000, 33, 18, 0.2), (6, 2500, 11, 11, 0.4),
                                                 (7, 900, 4, 4, 0.0), (8, 4000, 30, 30, 0.6)])
def test_quality_matches_reference_evaluation(tmp_path, seed, n, Bt, Bg, noise):
    """Block counts and the optimal matching total against the reference on every table;
    pairwise precision / recall and NMI on square tables only.  After the matching, the
    reference fills in the unmatched columns with a loop bounded by the row count
    (evaluate.hpp:87-97), which for a non-square table duplicates some columns and drops
    others before its pair counts and NMI are taken; on a square table every column is
    matched and its reordering is a permutation, which leaves those metrics unchanged."""
    rng = np.random.default_rng(seed)
    t = labels(rng, n, Bt)
    g = (t.astype(np.int64) * 7 + 3) % Bg
    flip = rng.random(n) < noise
    g[flip] = rng.integers(0, Bg, flip.sum())
    g[:Bg] = np.arange(Bg)
    g = g.astype(np.int32)
    q = gsap_amd.quality(t, g)
    r = ref_eval(tmp_path, t, g)
    assert q["B_truth"] == r["Number of partitions (truth)"]
    assert q["B_given"] == r["Number of partitions (given)"]
    # the optimal matching total is an integer: exact
    assert round(q["accuracy"] * n) == r["Matched total"]
    if Bt == Bg:
        # the reference accumulates in float: relative agreement to float precision
        assert q["recall"] == pytest.approx(r["Pairwise Recall"], rel=1e-5)
        assert q["precision"] == pytest.approx(r["Pairwise Precision"], rel=1e-5)
        assert q["nmi"] == pytest.approx(r["Normalized Mutual Information(NMI)"], rel=2e-4,
                                         abs=1e-5)


def test_accuracy_bruteforce_small_tables():
    rng = np.random.default_rng(9)
    for _ in range(60):
        n = int(rng.integers(5, 60))
        Bt, Bg = int(rng.integers(1, 6)), int(rng.integers(1, 6))
        t = rng.integers(0, Bt, n).astype(np.int32)
        g = rng.integers(0, Bg, n).astype(np.int32)
        T = np.zeros((t.max() + 1, g.max() + 1), np.int64)
        np.add.at(T, (t, g), 1)
        R, C = T.shape
        best = 0
        if R <= C:
            for perm in itertools.permutations(range(C), R):
                best = max(best, sum(T[i, perm[i]] for i in range(R)))
        else:
            for perm in itertools.permutations(range(R), C):
                best = max(best, sum(T[perm[j], j] for j in range(C)))
        assert round(gsap_amd.quality(t, g)["accuracy"] * n) == best


def test_quality_identities():
    rng = np.random.default_rng(4)
    t = rng.integers(0, 9, 5000).astype(np.int32)
    q = gsap_amd.quality(t, t)
    assert q["accuracy"] == 1.0 and q["precision"] == 1.0 and q["recall"] == 1.0
    assert q["rand_index"] == 1.0 and q["adjusted_rand_index"] == pytest.approx(1.0)
    assert q["nmi"] == pytest.approx(1.0, rel=1e-12)
    # a relabelling of the same partition scores the same
    perm = rng.permutation(9).astype(np.int32)
    q2 = gsap_amd.quality(t, perm[t])
    assert q2["accuracy"] == 1.0 and q2["nmi"] == pytest.approx(1.0, rel=1e-12)
    # pairwise precision / recall agree with the existing pair-count entry point
    g = rng.integers(0, 5, 5000).astype(np.int32)
    p, r = gsap_amd.pairwise(t, g)
    q3 = gsap_amd.quality(t, g)
    assert q3["precision"] == p and q3["recall"] == r
    # unknown truth (-1) is left out
    t4 = t.copy(); t4[:100] = -1
    q4 = gsap_amd.quality(t4, t)
    assert q4["n"] == 4900 and q4["accuracy"] == 1.0
    # independent labellings: NMI and ARI near 0
    a = rng.integers(0, 10, 200000).astype(np.int32)
    b = rng.integers(0, 10, 200000).astype(np.int32)
    q5 = gsap_amd.quality(a, b)
    assert q5["nmi"] < 1e-3 and abs(q5["adjusted_rand_index"]) < 1e-3
