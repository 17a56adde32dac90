"""Pins of the oracle's algorithm variants (delta_mode exact, Metropolis-Hastings acceptance).

Each check is independent of the oracle's formulas: the exact delta-S against a full
recomputation of the description length before and after the move or merge; the Hastings
ratio against a dense numpy evaluation of the proposal probabilities on the labellings
before and after the move; those probabilities against the oracle sampler's empirical
proposal frequencies; and the acceptance rule against u < exp(-beta dS) H with u from the
vertex's own Philox word."""
import math

import numpy as np
import pytest

import oracle_lib as O
from conftest import random_graph
from gsap_amd import ACCEPT_MH, DELTA_EXACT, Graph, Options

TAG_NODE = 0x4E4F4445


def dense_M(g, lab, B):
    M = np.zeros((B, B), np.int64)
    src = np.repeat(np.arange(g.N), np.diff(g.out_ptr))
    np.add.at(M, (lab[src], lab[g.out_adj]), g.out_w)
    return M


def graph_with_loops(seed, N, E, wmax):
    """Random graph whose vertices 0, 3, 6, ... carry a self loop."""
    g = random_graph(N, E, seed, wmax=wmax)
    src = list(np.repeat(np.arange(g.N), np.diff(g.out_ptr)))
    dst = list(g.out_adj)
    w = list(g.out_w)
    rng = np.random.default_rng(seed + 1)
    for v in range(0, N, 3):
        src.append(v); dst.append(v); w.append(int(rng.integers(1, wmax + 1)))
    return Graph.from_edges(N, src, dst, np.array(w, np.int32))


CASES = [(41, 30, 160, 4, 1), (42, 60, 420, 9, 3), (43, 45, 300, 20, 2)]


@pytest.mark.parametrize("seed,N,E,B,wmax", CASES)
def test_exact_nodal_delta_is_change_of_S(seed, N, E, B, wmax):
    g = graph_with_loops(seed, N, E, wmax)
    rng = np.random.default_rng(seed)
    lab = rng.integers(0, B, N).astype(np.int32)
    lab[0] = B - 1
    lab[lab == 1] = 2  # block 1 empty: moves into an empty block
    S0 = O.mdl(g, lab, B)
    worst_ref = 0.0
    for v in range(0, N, 2):
        for s in range(B):
            if s == lab[v]:
                continue
            d, _ = O.nodal_delta(g, lab, B, v, s, exact=1)
            l2 = lab.copy(); l2[v] = s
            true = O.mdl(g, l2, B) - S0
            assert abs(d - true) <= 1e-9 * (1 + abs(S0)), (v, s, d, true)
            dref, _ = O.nodal_delta(g, lab, B, v, s, exact=0)
            worst_ref = max(worst_ref, abs(dref - true))
    # the reference's delta differs (old column degrees in new rows, self loops)
    assert worst_ref > 1e-3


@pytest.mark.parametrize("seed,N,E,B,wmax", CASES)
def test_exact_merge_delta_is_change_of_S(seed, N, E, B, wmax):
    g = graph_with_loops(seed, N, E, wmax)
    lab = np.random.default_rng(seed + 7).integers(0, B, N).astype(np.int32)
    S0 = O.mdl(g, lab, B)
    worst_ref = 0.0
    for r in range(B):
        assert O.merge_delta(g, lab, B, r, r, exact=1) == np.finfo(np.float64).max
        for s in range(B):
            if s == r:
                continue
            d = O.merge_delta(g, lab, B, r, s, exact=1)
            l2 = np.where(lab == r, s, lab).astype(np.int32)
            true = O.mdl(g, l2, B) - S0  # same B: the model term cancels
            assert abs(d - true) <= 1e-9 * (1 + abs(S0)), (r, s, d, true)
            assert d >= -1e-9 * (1 + abs(S0))  # a merge never shortens the data part
            worst_ref = max(worst_ref, abs(O.merge_delta(g, lab, B, r, s, exact=0) - true))
    assert worst_ref > 1e-3


def proposal_prob(g, lab, B, v, target):
    """p(target | v) of the nodal sampler, evaluated densely on the labelling `lab`:
    neighbour entry j with probability w_j / k_v, then a uniform block with probability
    B / (d_t + B), else a block of t's row and column weighted by M."""
    M = dense_M(g, lab, B)
    d = M.sum(1) + M.sum(0)
    nb = np.concatenate([g.out_adj[g.out_ptr[v]:g.out_ptr[v + 1]],
                         g.in_adj[g.in_ptr[v]:g.in_ptr[v + 1]]])
    w = np.concatenate([g.out_w[g.out_ptr[v]:g.out_ptr[v + 1]],
                        g.in_w[g.in_ptr[v]:g.in_ptr[v + 1]]]).astype(np.float64)
    t = lab[nb]
    return float(np.sum(w / w.sum() * (M[t, target] + M[target, t] + 1) / (d[t] + B)))


@pytest.mark.parametrize("seed,N,E,B,wmax", CASES)
@pytest.mark.parametrize("exact", [0, 1])
def test_hastings_ratio_vs_dense(seed, N, E, B, wmax, exact):
    g = graph_with_loops(seed, N, E, wmax)
    lab = np.random.default_rng(seed + 3).integers(0, B, N).astype(np.int32)
    checked = loops = 0  # loops: checked moves of a vertex with a self loop
    for v in range(N):
        r = int(lab[v])
        for s in range(B):
            if s == r:
                continue
            _, lh = O.nodal_delta(g, lab, B, v, s, exact=exact)
            l2 = lab.copy(); l2[v] = s
            fwd = proposal_prob(g, lab, B, v, s)
            rev = proposal_prob(g, l2, B, v, r)
            assert abs(lh - (math.log2(rev) - math.log2(fwd))) < 1e-10, (v, s)
            checked += 1
            loops += int(np.any(g.out_adj[g.out_ptr[v]:g.out_ptr[v + 1]] == v))
    assert checked >= 50 and loops >= 10


def test_nodal_proposal_frequencies_match_closed_form():
    """The closed form the Hastings ratio uses is the distribution the sampler draws from
    (per-vertex frequencies over many batch counters, 5-sigma binomial bands)."""
    g = graph_with_loops(51, 40, 200, 3)
    B = 5
    lab = (np.arange(40) % B).astype(np.int32)
    o = Options(seed=17, num_batches=4)
    BS = 10
    trials = 3000
    counts = np.zeros((BS, B))
    for q in range(trials):
        res = O.nodal_batch(g, o, lab, B, 1, q)
        counts[np.arange(BS), res["prop"]] += 1
    for i in range(BS):
        v = BS + i
        p = np.array([proposal_prob(g, lab, B, v, s) for s in range(B)])
        assert abs(p.sum() - 1) < 1e-12
        sd = np.sqrt(trials * p * (1 - p)) + 1e-9
        assert np.all(np.abs(counts[i] - trials * p) <= 5 * sd + 1), (v, counts[i], trials * p)


@pytest.mark.parametrize("beta,exact", [(3.0, 1), (0.0, 0), (0.05, 1)])
def test_mh_decisions(beta, exact):
    g = graph_with_loops(61, 80, 500, 2)
    B = 7
    lab = np.random.default_rng(2).integers(0, B, 80).astype(np.int32)
    o = Options(seed=23, accept=ACCEPT_MH, beta=beta, delta_mode=exact, chain=4)
    b, q = 2, 11
    res = O.nodal_batch(g, o, lab, B, b, q)
    greedy = O.nodal_batch(g, Options(seed=23, delta_mode=exact, chain=4), lab, B, b, q)
    BS = 80 // 4
    acc = 0
    for i in range(BS):
        v = b * BS + i
        s = int(greedy["prop"][i])  # the proposal (MH's prop holds the final label)
        r = int(lab[v])
        if s == r:
            assert res["prop"][i] == r and res["dS"][i] == 0.0
            continue
        d, lh = O.nodal_delta(g, lab, B, v, s, exact=exact)
        x = O.philox([v, q, 4, TAG_NODE], [23, 0])[2]
        u = (x + 1) / 2.0 ** 32
        a = math.exp(-beta * d) * 2.0 ** lh
        if abs(u - a) < 1e-9 * a:
            continue  # too close to call in numpy
        want = u < a
        assert (res["prop"][i] == s) == want, (v, u, a)
        assert res["labels"][v] == (s if want else r)
        assert res["dS"][i] == (d if want else 0.0)
        acc += want
    assert sum(res["labels"][b * BS:(b + 1) * BS] != lab[b * BS:(b + 1) * BS]) == acc
    assert acc > 0


def test_mh_partition_runs_and_is_deterministic():
    g = graph_with_loops(71, 200, 1500, 1)
    o = Options(seed=5, accept=ACCEPT_MH, delta_mode=DELTA_EXACT, beta=3.0)
    lab, rep, tr = O.partition(g, o, trace_cap=4000)
    lab2, rep2, _ = O.partition(g, o)
    assert np.array_equal(lab, lab2) and rep.S == rep2.S
    assert 1 <= rep.B < g.N and rep.S == O.mdl(g, lab, rep.B)
    # some accepted moves raise S (MH), unlike the greedy rule
    assert any(t[0] == 2 and t[6] > 0 for t in tr) or any(t[0] == 2 and t[7] > 0 for t in tr)


def test_partition_without_edges():
    """No edges: one block, S = 0 (the model term would divide by E)."""
    for N in (1, 7):
        lab, rep, tr = O.partition(Graph.from_edges(N, [], []), Options(seed=1), trace_cap=10)
        assert rep.B == 1 and rep.S == 0.0 and rep.n_trace == 0
        assert np.array_equal(lab, np.zeros(N, np.int32))
