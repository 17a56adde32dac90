"""GPU parity of the algorithm variants: exact delta-S (nodal moves and merges) and
Metropolis-Hastings acceptance of nodal moves, bit-exact against the CPU oracle (whose
variants are pinned in test_oracle_variants.py).  Graphs with self loops exercise the
self-loop terms in every nodal form (register groups, one wave, a workgroup with LDS or
global-pool tables)."""
import numpy as np
import pytest

import gsap_amd
import oracle_lib as O
from conftest import hub_graph, random_graph, tsv_graph, with_self_loops
from gsap_amd import ACCEPT_MH, DELTA_EXACT, Options

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = gsap_amd.Context(0)
    yield c
    c.close()


CASES = {
    "loops_rand": lambda: with_self_loops(random_graph(700, 6000, 2, wmax=3), seed=1),
    "loops_hubs": lambda: with_self_loops(hub_graph(), seed=2),
    "loops_bighubs": lambda: with_self_loops(hub_graph(6, N=4000, hub_deg=1500), seed=3),
    "LL1k": lambda: tsv_graph("LL"),
    "HH5k": lambda: tsv_graph("HH", "5k"),
    # degrees spread over every nodal class (max(out, in) up to 16, 32, 64, 128 and above)
    "dcsbm3k": lambda: with_self_loops(gsap_amd.generate_dcsbm(3000, 8, 1.0, 2.0, 22.5, seed=9),
                                       seed=4),
}
_graphs = {}


def graph(name):
    if name not in _graphs:
        _graphs[name] = CASES[name]()
    return _graphs[name]


def eq_bits(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64),
                          np.asarray(b, np.float64).view(np.uint64))


def labels_for(g, B, seed):
    return np.random.default_rng(seed).integers(0, B, g.N).astype(np.int32)


VARIANTS = {
    "exact": dict(delta_mode=DELTA_EXACT),
    "mh": dict(accept=ACCEPT_MH, beta=3.0),
    "exact_mh": dict(delta_mode=DELTA_EXACT, accept=ACCEPT_MH, beta=3.0),
    "exact_mh_hot": dict(delta_mode=DELTA_EXACT, accept=ACCEPT_MH, beta=0.02),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("var", list(VARIANTS))
@pytest.mark.parametrize("Bfrac,dense", [(1.0, 0), (0.15, 0), (0.02, 0), (0.02, 1)])
def test_nodal_batch_variant(ctx, name, var, Bfrac, dense):
    g = graph(name)
    B = max(2, int(g.N * Bfrac))
    lab = labels_for(g, B, 8) if B < g.N else np.arange(g.N, dtype=np.int32)
    o = Options(seed=23, dense_max_B=dense, chain=2, **VARIANTS[var])
    moved = 0
    for b in range(o.num_batches):
        print(f"nodal {var} {name} B={B} dense={dense} batch={b}", flush=True)
        got = ctx.nodal_batch(g, o, lab, B, b, 11 + b)
        ref = O.nodal_batch(g, o, lab, B, b, 11 + b)
        assert np.array_equal(got["prop"], ref["prop"])
        assert eq_bits(got["dS"], ref["dS"])
        assert eq_bits(got["dE"], ref["dE"])
        assert np.array_equal(got["labels"], ref["labels"])
        moved += int(np.sum(got["labels"] != lab))
    # from singleton blocks (B = N) a move merges two blocks, which never shortens the data
    # part of S: with the exact delta no move is accepted greedily, and few or none under MH
    if B < g.N:
        assert moved > 0


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("Bfrac,dense", [(1.0, 0), (0.2, 0), (0.03, 0), (0.03, 1)])
def test_merge_phase_exact(ctx, name, Bfrac, dense):
    g = graph(name)
    B = max(2, int(g.N * Bfrac))
    lab = labels_for(g, B, 6) if B < g.N else np.arange(g.N, dtype=np.int32)
    o = Options(seed=17, dense_max_B=dense, delta_mode=DELTA_EXACT)
    print(f"merge exact {name} B={B} dense={dense}", flush=True)
    got = ctx.merge_phase(g, o, lab, B, int(B * 0.35), 3)
    ref = O.merge_phase(g, o, lab, B, int(B * 0.35), 3)
    assert np.array_equal(got["prop"], ref["prop"])
    assert eq_bits(got["dS"], ref["dS"])
    assert np.array_equal(got["best"], ref["best"])
    assert eq_bits(got["dS_best"], ref["dS_best"])
    assert got["B"] == ref["B"]
    assert np.array_equal(got["labels"], ref["labels"])


@pytest.mark.parametrize("name", ["loops_hubs", "LL1k", "dcsbm3k"])
@pytest.mark.parametrize("var", ["exact", "mh", "exact_mh"])
def test_partition_trace_variant(ctx, name, var):
    g = graph(name)
    o = Options(seed=3, trace=1, **VARIANTS[var])
    print(f"partition {var} {name}", flush=True)
    lab, rep = ctx.partition(g, o)
    tr = ctx.trace()
    rlab, rrep, rtr = O.partition(g, o, trace_cap=len(tr) + 10)
    assert len(tr) == rrep.n_trace
    for k, (a, b) in enumerate(zip(tr, rtr)):
        assert a[:6] == b[:6], (k, a, b)
        assert eq_bits(a[6], b[6]) and eq_bits(a[7], b[7]), (k, a, b)
    assert np.array_equal(lab, rlab)
    assert rep.B == rrep.B and eq_bits(rep.S, rrep.S)


@pytest.mark.parametrize("name", ["loops_rand", "loops_hubs", "HH5k", "dcsbm3k"])
@pytest.mark.parametrize("Bfrac", [0.15, 0.02])
def test_nodal_batch_mixed_exact(ctx, name, Bfrac):
    """Mixed form with the exact delta: FP32 rejections, FP64 for the rest; same bits as the
    FP64 oracle."""
    g = graph(name)
    B = max(2, int(g.N * Bfrac))
    lab = labels_for(g, B, 9)
    o = Options(seed=29, precision=gsap_amd.PRECISION_MIXED, delta_mode=DELTA_EXACT)
    for b in range(o.num_batches):
        print(f"nodal mixed exact {name} B={B} batch={b}", flush=True)
        got = ctx.nodal_batch(g, o, lab, B, b, 5 + b)
        ref = O.nodal_batch(g, o, lab, B, b, 5 + b)
        assert np.array_equal(got["prop"], ref["prop"])
        assert eq_bits(got["dS"], ref["dS"])
        assert eq_bits(got["dE"], ref["dE"])


def test_variant_options_checked(ctx):
    g = graph("LL1k")
    with pytest.raises(gsap_amd.GsapError, match="needs precision FP64"):
        ctx.partition(g, Options(accept=ACCEPT_MH, precision=gsap_amd.PRECISION_MIXED))
    with pytest.raises(gsap_amd.GsapError, match="beta"):
        ctx.partition(g, Options(accept=ACCEPT_MH, beta=-1.0))
    with pytest.raises(gsap_amd.GsapError, match="delta_mode"):
        ctx.partition(g, Options(delta_mode=7))


def _edge_cases():
    rng = np.random.default_rng(5)
    star_src = [0] * 40 + list(range(1, 41))
    star_dst = list(range(1, 41)) + [0] * 40
    clique = [(a, b) for a in range(12) for b in range(12) if a != b]
    iso_src = rng.integers(0, 50, 120)
    iso_dst = (iso_src + 1 + rng.integers(0, 48, 120)) % 50
    pairs = sorted(set(zip(iso_src.tolist(), iso_dst.tolist())))
    return {
        "one_vertex": gsap_amd.Graph.from_edges(1, [], []),
        "no_edges": gsap_amd.Graph.from_edges(7, [], []),
        "isolated": gsap_amd.Graph.from_edges(90, [p[0] for p in pairs], [p[1] for p in pairs]),
        "star": gsap_amd.Graph.from_edges(41, star_src, star_dst),
        "clique": gsap_amd.Graph.from_edges(12, [a for a, _ in clique], [b for _, b in clique]),
        "self_loops_only": gsap_amd.Graph.from_edges(9, list(range(9)), list(range(9)),
                                                     np.arange(1, 10, dtype=np.int32)),
        "odd_N": with_self_loops(random_graph(203, 900, 8, wmax=5), step=7, seed=4),
    }


EDGE = _edge_cases()


@pytest.mark.parametrize("name", list(EDGE))
@pytest.mark.parametrize("var", ["reference", "exact_mh"])
def test_partition_edge_cases(ctx, name, var):
    """Degenerate inputs: a single vertex, no edges, isolated vertices, N not a multiple of
    the batch count (its tail never moves), stars, cliques, self loops only; the whole
    trace equals the oracle's."""
    g = EDGE[name]
    o = Options(seed=7, trace=1, **({} if var == "reference" else VARIANTS[var]))
    print(f"edge case {name} {var}", flush=True)
    lab, rep = ctx.partition(g, o)
    tr = ctx.trace()
    rlab, rrep, rtr = O.partition(g, o, trace_cap=len(tr) + 10)
    assert len(tr) == rrep.n_trace
    for k, (a, b) in enumerate(zip(tr, rtr)):
        assert a[:6] == b[:6], (k, a, b)
        assert eq_bits(a[6], b[6]) and eq_bits(a[7], b[7]), (k, a, b)
    assert np.array_equal(lab, rlab) and rep.B == rrep.B and eq_bits(rep.S, rrep.S)
    assert lab.min() >= 0 and lab.max() < max(rep.B, 1) or g.N == 1
