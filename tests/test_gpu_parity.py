"""GPU parity: the HIP path against the CPU oracle, bit-exact.

Integer outputs (block matrix, proposals, labels, merges) must be identical; FP64 delta-S,
batch sums and description lengths must be identical bit for bit (same canonical log2 and
summation order on both sides)."""
import numpy as np
import pytest

import gsap_amd
import oracle_lib as O
from conftest import hub_graph, random_graph, tsv_graph
from gsap_amd import Options

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = gsap_amd.Context(0)
    yield c
    c.close()


CASES = {
    "tiny_w": lambda: random_graph(64, 400, 1, wmax=4),
    "rand": lambda: random_graph(700, 6000, 2),
    "hubs": hub_graph,
    "bighubs": lambda: hub_graph(6, N=4000, hub_deg=1500),
    "LL1k": lambda: tsv_graph("LL"),
    "HH5k": lambda: tsv_graph("HH", "5k"),
    # degrees spread over every nodal class (max(out, in) up to 16, 32, 64, 128 and above)
    "dcsbm3k": lambda: gsap_amd.generate_dcsbm(3000, 8, 1.0, 2.0, 22.5, seed=9),
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
    lab = np.random.default_rng(seed).integers(0, B, g.N).astype(np.int32)
    return lab


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("Bfrac,dense", [(1.0, 0), (0.25, 0), (0.01, 0), (0.01, 1), (0.25, 100000)])
def test_blockmodel(ctx, name, Bfrac, dense):
    g = graph(name)
    B = max(1, int(g.N * Bfrac))
    lab = labels_for(g, B, 3) if B < g.N else np.arange(g.N, dtype=np.int32)
    print(f"blockmodel {name} B={B} dense_max_B={dense}", flush=True)
    got = ctx.blockmodel(g, lab, B, Options(dense_max_B=dense))
    ref = O.blockmodel(g, lab, B)
    for k in ("nnz", "rowptr", "col", "val", "colptr", "row", "cval", "dout", "din"):
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("Bfrac", [1.0, 0.3, 0.02])
def test_mdl(ctx, name, Bfrac):
    g = graph(name)
    B = max(1, int(g.N * Bfrac))
    lab = labels_for(g, B, 4) if B < g.N else np.arange(g.N, dtype=np.int32)
    assert eq_bits(ctx.mdl(g, lab, B), O.mdl(g, lab, B))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("Bfrac,dense", [(1.0, 0), (0.2, 0), (0.03, 0), (0.03, 1)])
def test_merge_phase(ctx, name, Bfrac, dense):
    g = graph(name)
    B = max(2, int(g.N * Bfrac))
    lab = labels_for(g, B, 6) if B < g.N else np.arange(g.N, dtype=np.int32)
    o = Options(seed=17, dense_max_B=dense)
    print(f"merge {name} B={B}", flush=True)
    got = ctx.merge_phase(g, o, lab, B, int(B * 0.35), 3)
    ref = O.merge_phase(g, o, lab, B, int(B * 0.35), 3)
    assert np.array_equal(got["prop"], ref["prop"])
    assert eq_bits(got["dS"], ref["dS"])
    assert np.array_equal(got["best"], ref["best"])
    assert eq_bits(got["dS_best"], ref["dS_best"])
    assert got["B"] == ref["B"]
    assert np.array_equal(got["labels"], ref["labels"])


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("Bfrac,dense", [(1.0, 0), (0.15, 0), (0.02, 0), (0.02, 1)])
def test_nodal_batch(ctx, name, Bfrac, dense):
    g = graph(name)
    B = max(2, int(g.N * Bfrac))
    lab = labels_for(g, B, 8) if B < g.N else np.arange(g.N, dtype=np.int32)
    o = Options(seed=23, dense_max_B=dense)
    for b in range(o.num_batches):
        print(f"nodal {name} B={B} batch={b}", flush=True)
        got = ctx.nodal_batch(g, o, lab, B, b, 11 + b)
        ref = O.nodal_batch(g, o, lab, B, b, 11 + b)
        assert np.array_equal(got["prop"], ref["prop"])
        assert eq_bits(got["dS"], ref["dS"])
        assert eq_bits(got["dE"], ref["dE"])
        assert np.array_equal(got["labels"], ref["labels"])


@pytest.mark.parametrize("name,model,csr", [("tiny_w", 0, 0), ("hubs", 0, 0), ("LL1k", 0, 0),
                                            ("LL1k", 1, 0), ("HH5k", 1, 0), ("tiny_w", 0, 1),
                                            ("HH5k", 1, 1)])
def test_partition_trace(ctx, name, model, csr):
    """Whole traces against the oracle: by default the sweeps' batches before the last
    update only the dense block matrix and the block degrees (samples walk the dense
    matrix); csr = 1 rebuilds the CSR/CSC after every batch."""
    g = graph(name)
    o = Options(seed=3, trace=1, model=model, csr_every_batch=csr)
    print(f"partition {name} csr_every_batch={csr}", flush=True)
    lab, rep = ctx.partition(g, o)
    tr = ctx.trace()
    rlab, rrep, rtr = O.partition(g, o, trace_cap=len(tr) + 10)
    assert len(tr) == rrep.n_trace
    for k, (a, b) in enumerate(zip(tr, rtr)):
        assert a[:6] == b[:6], (k, a, b)
        assert eq_bits(a[6], b[6]) and eq_bits(a[7], b[7]), (k, a, b)
    assert np.array_equal(lab, rlab)
    assert rep.B == rrep.B and eq_bits(rep.S, rrep.S)
    assert eq_bits(rep.S, ctx.mdl(g, lab, rep.B, Options(model=model)))


def weighted_dcsbm():
    """20000 vertices, 120k edges of weight 1..3: the search passes B = 13000 and 8450 (dense
    block matrix with the CSR deferred inside sweeps and column chunk sums) and the dense
    ranges below with weighted degrees."""
    g = gsap_amd.generate_dcsbm(20000, 20, 1.0, 2.0, 6.0, seed=13)
    src = np.repeat(np.arange(g.N), np.diff(g.out_ptr))
    w = np.random.default_rng(13).integers(1, 4, g.E).astype(np.int32)
    return gsap_amd.Graph.from_edges(g.N, src, g.out_adj, w)


def test_partition_trace_weighted_dense_range(ctx):
    g = weighted_dcsbm()
    o = Options(seed=7, trace=1)
    print("partition weighted 20k", flush=True)
    lab, rep = ctx.partition(g, o)
    tr = ctx.trace()
    assert any(r[0] == 2 and 8192 < r[1] <= 16384 for r in tr)
    rlab, rrep, rtr = O.partition(g, o, trace_cap=len(tr) + 10, threads=8)
    assert len(tr) == rrep.n_trace
    for k, (a, b) in enumerate(zip(tr, rtr)):
        assert a[:6] == b[:6], (k, a, b)
        assert eq_bits(a[6], b[6]) and eq_bits(a[7], b[7]), (k, a, b)
    assert np.array_equal(lab, rlab)
    assert rep.B == rrep.B and eq_bits(rep.S, rrep.S)


def test_partition_generated_c1(ctx):
    g = gsap_amd.generate_dcsbm(1000, 11, 0.0, 4.0, 20.0, seed=1)
    o = Options(seed=1)
    lab, rep = ctx.partition(g, o)
    rlab, rrep, _ = O.partition(g, o)
    assert np.array_equal(lab, rlab) and eq_bits(rep.S, rrep.S)
    assert rep.S == ctx.mdl(g, lab, rep.B)


def test_partition_generated_c2(ctx):
    g = gsap_amd.generate_dcsbm(20000, 33, 1.0, 1.5, 20.0, seed=2)
    o = Options(seed=1)
    print("partition C2", flush=True)
    lab, rep = ctx.partition(g, o)
    rlab, rrep, _ = O.partition(g, o)
    assert rep.B == rrep.B and eq_bits(rep.S, rrep.S)
    assert np.array_equal(lab, rlab)


@pytest.fixture(scope="module")
def c3():
    return gsap_amd.generate_dcsbm(200000, 71, 1.0, 1.5, 22.5, seed=3)


@pytest.mark.parametrize("B,dense", [(71, 0), (5000, 0), (200000, 0), (20000, 0), (3000, 1)])
def test_c3_blockmodel_and_batch(ctx, c3, B, dense):
    """dense = 1 (and B > 24576) look the block matrix up in CSR/CSC rows of hundreds of
    entries through the coarse index."""
    g = c3
    lab = labels_for(g, B, 9) if B < g.N else np.arange(g.N, dtype=np.int32)
    print(f"C3 blockmodel B={B} dense={dense}", flush=True)
    got = ctx.blockmodel(g, lab, B, Options(dense_max_B=dense))
    ref = O.blockmodel(g, lab, B)
    for k in ("nnz", "rowptr", "col", "val", "colptr", "row", "cval", "dout", "din"):
        assert np.array_equal(got[k], ref[k]), k
    # size-independent invariants
    assert got["val"].sum() == g.E and got["dout"].sum() == g.E and got["din"].sum() == g.E
    o = Options(seed=5, dense_max_B=dense)
    print(f"C3 nodal B={B}", flush=True)
    gb = ctx.nodal_batch(g, o, lab, B, 1, 40)
    rb = O.nodal_batch(g, o, lab, B, 1, 40)
    assert np.array_equal(gb["prop"], rb["prop"]) and eq_bits(gb["dS"], rb["dS"])
    assert eq_bits(gb["dE"], rb["dE"])


def heavy_graph():
    """65536 vertices, 60000 edges of weight 1..4 and one of weight 300000: 19 weight bits,
    so the segmented rebuild packs (block, weight) into one word only while B < 8192."""
    g = random_graph(65536, 60000, 12, wmax=4)
    src = np.repeat(np.arange(g.N), np.diff(g.out_ptr))
    w = g.out_w.copy()
    w[7] = 300000
    return gsap_amd.Graph.from_edges(g.N, src, g.out_adj, w)


@pytest.fixture(scope="module")
def heavy():
    return heavy_graph()


@pytest.mark.parametrize("reg,cnt", [(16, 4096), (32, 1024), (64, 4096)])
@pytest.mark.parametrize("B,big", [(9000, 0.3), (3000, 0.05), (20000, 0.3)])
def test_blockmodel_segment_classes(ctx, heavy, B, big, reg, cnt):
    """The segment-sort classes only move segments between sorting teams: register groups
    up to seg_reg_max entries, and (B <= 16000) workgroup sorts instead of LDS counting up to
    seg_count_min entries; every choice gives the oracle's block matrix."""
    g = heavy
    rng = np.random.default_rng(17)
    lab = rng.integers(1, B, g.N).astype(np.int32)
    lab[rng.random(g.N) < big] = 0
    print(f"heavy blockmodel B={B} big={big} seg_reg_max={reg} seg_count_min={cnt}", flush=True)
    got = ctx.blockmodel(g, lab, B, Options(dense_max_B=1, tile_max_B=1, seg_reg_max=reg,
                                            seg_count_min=cnt))
    ref = O.blockmodel(g, lab, B)
    for k in ("nnz", "rowptr", "col", "val", "colptr", "row", "cval", "dout", "din"):
        assert np.array_equal(got[k], ref[k]), k


NODAL_CLASSES = [(-1, 32, 128), (16, -1, 128), (8, 32, 64), (-1, -1, -1), (16, 32, 96)]


@pytest.mark.parametrize("bounds", NODAL_CLASSES)
@pytest.mark.parametrize("name", ["dcsbm3k", "bighubs"])
def test_partition_nodal_classes(ctx, name, bounds):
    """The nodal neighbour-count classes only choose the team of a vertex (16 or 32 lanes,
    a wave, a workgroup; -1 leaves a form unused); the partition is the oracle's."""
    g = graph(name)
    m16, m32, mw = bounds
    o = Options(seed=6, nodal_max16=m16, nodal_max32=m32, nodal_maxwave=mw, seg_reg_max=32,
                seg_count_min=2048)
    print(f"partition {name} nodal classes {bounds}", flush=True)
    lab, rep = ctx.partition(g, o)
    rlab, rrep, _ = O.partition(g, Options(seed=6))
    assert np.array_equal(lab, rlab) and eq_bits(rep.S, rrep.S)


def test_rejects_bad_work_split(ctx):
    g = graph("rand")
    for kw in (dict(nodal_max16=17), dict(nodal_max32=33), dict(nodal_maxwave=129),
               dict(nodal_max16=-2), dict(seg_reg_max=48), dict(seg_count_min=100),
               dict(seg_count_min=5000)):
        with pytest.raises(gsap_amd.GsapError):
            ctx.partition(g, Options(seed=1, **kw))


@pytest.mark.parametrize("B,big", [(65536, 0.0), (20000, 0.3), (9000, 0.3), (3000, 0.05),
                                   (9000, 0.0)])
def test_blockmodel_unpacked_and_packed(ctx, heavy, B, big):
    """Segment entries as separate key and weight words (B >= 8192 here: sorts, one-range
    and two-range counting) and packed into one word (B < 8192), on a skewed labelling."""
    g = heavy
    rng = np.random.default_rng(13)
    lab = rng.integers(1, B, g.N).astype(np.int32) if B < g.N else np.arange(g.N, dtype=np.int32)
    lab[rng.random(g.N) < big] = 0
    print(f"heavy blockmodel B={B} big={big}", flush=True)
    got = ctx.blockmodel(g, lab, B, Options(dense_max_B=1, tile_max_B=1))
    ref = O.blockmodel(g, lab, B)
    for k in ("nnz", "rowptr", "col", "val", "colptr", "row", "cval", "dout", "din"):
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("B,big,dense", [(20000, 0.3, 0), (20000, 0.3, 30000), (3000, 0.05, 0),
                                         (300, 0.0, 0), (300, 0.0, 1), (129, 0.5, 0)])
def test_c3_blockmodel_skewed(ctx, c3, B, big, dense):
    """Labellings whose segments fall in every size class of the segmented rebuild: one
    block holding a fraction `big` of the vertices (segments of ~1M entries, counted over
    several column ranges when B > 16000), the rest uniform."""
    g = c3
    rng = np.random.default_rng(11)
    lab = rng.integers(1, B, g.N).astype(np.int32)
    lab[rng.random(g.N) < big] = 0
    print(f"C3 skewed blockmodel B={B} big={big} dense={dense}", flush=True)
    got = ctx.blockmodel(g, lab, B, Options(dense_max_B=dense))
    ref = O.blockmodel(g, lab, B)
    for k in ("nnz", "rowptr", "col", "val", "colptr", "row", "cval", "dout", "din"):
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("B,big,dense,exact", [(3000, 0.3, 0, 0), (3000, 0.0, 0, 1),
                                               (20000, 0.0, 0, 0), (900, 0.05, 1, 0)])
def test_c3_merge_phase(ctx, c3, B, big, dense, exact):
    """Merge phases with long rows (blocks staged in LDS, 512 or 2048 entries) and one block
    holding a fraction `big` of the vertices (its row and column exceed the staging arrays)."""
    g = c3
    rng = np.random.default_rng(21)
    lab = rng.integers(1, B, g.N).astype(np.int32)
    lab[rng.random(g.N) < big] = 0
    o = Options(seed=31, dense_max_B=dense, delta_mode=exact)
    print(f"C3 merge B={B} big={big} dense={dense} exact={exact}", flush=True)
    got = ctx.merge_phase(g, o, lab, B, int(B * 0.35), 7)
    ref = O.merge_phase(g, o, lab, B, int(B * 0.35), 7)
    assert np.array_equal(got["prop"], ref["prop"])
    assert eq_bits(got["dS"], ref["dS"])
    assert np.array_equal(got["best"], ref["best"])
    assert got["B"] == ref["B"] and np.array_equal(got["labels"], ref["labels"])


BM_KEYS = ("nnz", "rowptr", "col", "val", "colptr", "row", "cval", "dout", "din")


@pytest.mark.parametrize("name,B,big,dense", [("c3", 20000, 0.3, 0), ("c3", 20000, 0.3, 1),
                                              ("c3", 3000, 0.05, 1),
                                              ("c3", 200000, 0.0, 0), ("c3", 9000, 0.0, 0),
                                              ("heavy", 9000, 0.3, 1), ("heavy", 20000, 0.0, 1),
                                              ("hubs", 150, 0.0, 1), ("tiny_w", 16, 0.0, 1),
                                              ("bighubs", 1000, 0.1, 1)])
def test_block_matrix_after_batch(ctx, c3, heavy, name, B, big, dense):
    """The block matrix the partition loop holds after each nodal batch of a sweep (without
    the dense copy: the batch's moves merged into the current CSR/CSC as a sorted signed delta
    per row and column; B > 16000 with one block of 30 % of the vertices puts delta segments
    of more than 4096 entries on the column-range counts) equals the oracle's block matrix
    of the batch's labels, and the full rebuild (sparse_update = 0) gives the same."""
    g = {"c3": c3, "heavy": heavy}.get(name) or graph(name)
    rng = np.random.default_rng(29)
    lab = rng.integers(1, B, g.N).astype(np.int32) if B < g.N else np.arange(g.N, dtype=np.int32)
    if big:
        lab[rng.random(g.N) < big] = 0
    o = Options(seed=41, dense_max_B=dense, sparse_update=1)
    for b in range(o.num_batches):
        print(f"batch and block matrix {name} B={B} big={big} dense={dense} batch={b}", flush=True)
        got_lab, got = ctx.nodal_batch_bm(g, o, lab, B, b, 60 + b)
        ref_lab = O.nodal_batch(g, o, lab, B, b, 60 + b)["labels"]
        assert np.array_equal(got_lab, ref_lab)
        ref = O.blockmodel(g, got_lab, B)
        for k in BM_KEYS:
            assert np.array_equal(got[k], ref[k]), (b, k)
        full_lab, full = ctx.nodal_batch_bm(g, Options(seed=41, dense_max_B=dense), lab, B, b,
                                            60 + b)
        assert np.array_equal(full_lab, got_lab)
        for k in BM_KEYS:
            assert np.array_equal(full[k], got[k]), (b, k)
        lab = got_lab


def test_block_matrix_consecutive_batches(ctx, c3):
    """Two sweeps of batches after the first merge phase of the 200k-vertex graph (B =
    130000), each continuing from the block matrix the previous batch's sparse update left
    (rows of hundreds of entries next to delta rows of a few: chunks where 64 current entries
    precede the next delta run): every step equals the full rebuild of the same labels, and
    the last one the oracle's block matrix."""
    g = c3
    o = Options(seed=1, sparse_update=1)
    m = ctx.merge_phase(g, o, np.arange(g.N, dtype=np.int32), g.N, int(g.N * 0.35), 0)
    lab, B = m["labels"], m["B"]
    with gsap_amd.Context(0) as full_ctx:
        for q in range(8):
            print(f"consecutive batch {q} B={B}", flush=True)
            lab, got = ctx.nodal_batch_bm(g, o, lab, B, q % 4, q, keep=q > 0)
            full = full_ctx.blockmodel(g, lab, B)
            for k in BM_KEYS:
                assert np.array_equal(got[k], full[k]), (q, k)
    ref = O.blockmodel(g, lab, B)
    for k in BM_KEYS:
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("name,dense", [("hubs", 1), ("LL1k", 64), ("HH5k", 1),
                                        ("tiny_w", 1)])
def test_partition_trace_sparse_update(ctx, name, dense):
    """Whole partitions whose levels above the dense range update the block matrix after
    every batch by merging the moves: the trace (labels, proposals, block-matrix hashes per
    sweep) equals the oracle's."""
    g = graph(name)
    o = Options(seed=6, trace=1, dense_max_B=dense, sparse_update=1)
    print(f"partition sparse update {name} dense_max_B={dense}", flush=True)
    lab, rep = ctx.partition(g, o)
    tr = ctx.trace()
    rlab, rrep, rtr = O.partition(g, o, trace_cap=len(tr) + 10)
    assert len(tr) == rrep.n_trace
    for k, (a, b) in enumerate(zip(tr, rtr)):
        assert a[:6] == b[:6], (k, a, b)
        assert eq_bits(a[6], b[6]) and eq_bits(a[7], b[7]), (k, a, b)
    assert np.array_equal(lab, rlab) and eq_bits(rep.S, rrep.S)


@pytest.mark.parametrize("B", [300, 2500, 6000])
@pytest.mark.parametrize("tile", [1, 8192])
def test_c3_blockmodel_tile_threshold(ctx, c3, B, tile):
    """tile_max_B only picks the rebuild path (row tile up to 8192 blocks, else segmented):
    both give the oracle's block matrix."""
    g = c3
    lab = labels_for(g, B, 13)
    print(f"C3 blockmodel B={B} tile_max_B={tile}", flush=True)
    got = ctx.blockmodel(g, lab, B, Options(dense_max_B=8192, tile_max_B=tile))
    ref = O.blockmodel(g, lab, B)
    for k in ("nnz", "rowptr", "col", "val", "colptr", "row", "cval", "dout", "din"):
        assert np.array_equal(got[k], ref[k]), k


def test_autotune_keeps_results(ctx):
    """The autotuner measures the path thresholds and the work-split classes on snapshots of
    the caller's partition; the tuned options give the oracle's partition."""
    g = gsap_amd.generate_dcsbm(20000, 33, 1.0, 1.5, 20.0, seed=2)
    print("autotune C2", flush=True)
    o = ctx.autotune(g, Options(seed=1))
    assert 128 <= o.tile_max_B <= 8192 and o.dense_max_B >= 2047
    assert o.seg_reg_max in (0, 16, 32) and o.seg_count_min in (0, 4096)
    assert o.nodal_max16 in (0, -1, 8) and o.nodal_max32 in (0, -1)
    assert o.nodal_maxwave in (0, 64, 96)
    lab, rep = ctx.partition(g, o)
    rlab, rrep, _ = O.partition(g, Options(seed=1))
    assert np.array_equal(lab, rlab) and eq_bits(rep.S, rrep.S)


@pytest.mark.parametrize("name,dense,sparse", [("hubs", 0, 0), ("LL1k", 0, 0), ("HH5k", 0, 0),
                                               ("HH5k", 64, 0), ("HH5k", 64, 1)])
def test_partition_graph_replay(ctx, name, dense, sparse):
    """options.graph: every level's sweep captured once and replayed (at least three times
    per level, with the level's eager merge, rebuild and golden-section work in between);
    labels and S must equal the oracle's, and a second call must reproduce the first.  With
    dense_max_B = 64 the levels above 64 blocks capture the segmented rebuild and the CSR
    lookups instead of the dense block matrix (with sparse = 1, the sparse update after each
    batch, which alternates between two buffer sets)."""
    g = graph(name)
    o = Options(seed=4, graph=1, dense_max_B=dense, sparse_update=sparse)
    print(f"partition graph replay {name} dense_max_B={dense}", flush=True)
    lab, rep = ctx.partition(g, o)
    rlab, rrep, _ = O.partition(g, Options(seed=4))
    assert np.array_equal(lab, rlab) and eq_bits(rep.S, rrep.S)
    assert rep.sweeps == rrep.sweeps and rep.B == rrep.B
    lab2, rep2 = ctx.partition(g, o)
    assert np.array_equal(lab2, lab) and eq_bits(rep2.S, rep.S)


def test_rejects_malformed_graphs(ctx):
    """Host-side validation before any device work: an in-CSR that does not hold the
    out-CSR's edges (a moved or reweighted in-edge), and ids out of range."""
    g = graph("rand")
    bad = gsap_amd.Graph(g.N, g.out_ptr, g.out_adj, g.out_w, g.in_ptr, g.in_adj.copy(),
                         g.in_w)
    bad.in_adj[0] = (bad.in_adj[0] + 1) % g.N
    with pytest.raises(gsap_amd.GsapError, match="different edges"):
        ctx.partition(bad, Options(seed=1))
    bad = gsap_amd.Graph(g.N, g.out_ptr, g.out_adj, g.out_w, g.in_ptr, g.in_adj,
                         g.in_w.copy())
    bad.in_w[3] += 1
    with pytest.raises(gsap_amd.GsapError, match="different edges"):
        ctx.partition(bad, Options(seed=1))
    bad = gsap_amd.Graph(g.N, g.out_ptr, g.out_adj.copy(), g.out_w, g.in_ptr, g.in_adj, g.in_w)
    bad.out_adj[5] = g.N
    with pytest.raises(gsap_amd.GsapError, match="out of range"):
        ctx.partition(bad, Options(seed=1))
    # the context stays usable
    lab, rep = ctx.partition(g, Options(seed=1))
    rlab, rrep, _ = O.partition(g, Options(seed=1))
    assert np.array_equal(lab, rlab)


def test_partition_resident(ctx):
    """gsap_upload_graph_i32 + gsap_partition_resident_i32: the oracle's partition, twice in
    a row from the one upload; a call that takes another graph replaces the resident copy;
    a wrong N is an error and the context stays usable."""
    g, h = graph("LL1k"), graph("hubs")
    o = Options(seed=4)
    rlab, rrep, _ = O.partition(g, o)
    ctx.upload_graph(g)
    for _ in range(2):
        lab, rep = ctx.partition_resident(g.N, o)
        assert np.array_equal(lab, rlab) and eq_bits(rep.S, rrep.S) and rep.B == rrep.B
    with pytest.raises(gsap_amd.GsapError, match="resident graph"):
        ctx.partition_resident(g.N + 1, o)
    hlab, hrep = ctx.partition(h, o)
    lab, rep = ctx.partition_resident(h.N, o)
    assert np.array_equal(lab, hlab) and eq_bits(rep.S, hrep.S)


def test_partition_batch(ctx):
    graphs = [graph("tiny_w"), graph("hubs"), graph("LL1k"),
              gsap_amd.generate_dcsbm(1000, 11, 0.0, 4.0, 20.0, seed=1), graph("rand")]
    o = Options(seed=9, workers=3)
    outs, reps = ctx.partition_batch(graphs, o)
    for g, lab, rep in zip(graphs, outs, reps):
        rlab, rrep, _ = O.partition(g, o)
        assert np.array_equal(lab, rlab) and eq_bits(rep.S, rrep.S)


def test_partition_chains(ctx):
    g = graph("LL1k")
    o = Options(seed=2, chain=5, workers=2)
    lab, rep, S, best = ctx.partition_chains(g, 4, o)
    refs = [O.partition(g, Options(seed=2, chain=5 + k)) for k in range(4)]
    for k in range(4):
        assert eq_bits(S[k], refs[k][1].S)
    assert best == int(np.argmin([r[1].S for r in refs]))
    assert np.array_equal(lab, refs[best][0]) and eq_bits(rep.S, S[best])


MIXED = gsap_amd.PRECISION_MIXED


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("Bfrac,dense", [(1.0, 0), (0.15, 0), (0.02, 0), (0.02, 1)])
def test_nodal_batch_mixed(ctx, name, Bfrac, dense):
    """Mixed form (FP32 delta-S, FP64 for accepted and near-threshold proposals): same
    proposals, delta-S bits, batch sum and labels as the FP64 oracle."""
    g = graph(name)
    B = max(2, int(g.N * Bfrac))
    lab = labels_for(g, B, 8) if B < g.N else np.arange(g.N, dtype=np.int32)
    o = Options(seed=23, dense_max_B=dense, precision=MIXED)
    for b in range(o.num_batches):
        print(f"nodal mixed {name} B={B} batch={b}", flush=True)
        got = ctx.nodal_batch(g, o, lab, B, b, 11 + b)
        ref = O.nodal_batch(g, o, lab, B, b, 11 + b)
        assert np.array_equal(got["prop"], ref["prop"])
        assert eq_bits(got["dS"], ref["dS"])
        assert eq_bits(got["dE"], ref["dE"])
        assert np.array_equal(got["labels"], ref["labels"])


@pytest.mark.parametrize("name", ["hubs", "LL1k", "HH5k"])
def test_partition_trace_mixed(ctx, name):
    g = graph(name)
    o = Options(seed=3, trace=1, precision=MIXED)
    print(f"partition mixed {name}", flush=True)
    lab, rep = ctx.partition(g, o)
    tr = ctx.trace()
    rlab, rrep, rtr = O.partition(g, o, trace_cap=len(tr) + 10)
    assert len(tr) == rrep.n_trace
    for k, (a, b) in enumerate(zip(tr, rtr)):
        assert a[:6] == b[:6], (k, a, b)
        assert eq_bits(a[6], b[6]) and eq_bits(a[7], b[7]), (k, a, b)
    assert np.array_equal(lab, rlab) and eq_bits(rep.S, rrep.S)
    # every accepted move was evaluated in FP64; the FP32 error is tiny next to S
    assert 0 < rep.fp64_evals <= rep.batches * (g.N // 4)
    assert 0 <= rep.rechecks <= rep.fp64_evals
    assert 0 < rep.max_abs_err_fp32 < 1e-6 * abs(rep.S)


def test_partition_c3_mixed_equals_fp64(ctx, c3):
    """200k-vertex graph: the mixed form partitions bit-identically to the plain FP64 form."""
    print("C3 fp64", flush=True)
    lab64, rep64 = ctx.partition(c3, Options(seed=1))
    print("C3 mixed", flush=True)
    lab32, rep32 = ctx.partition(c3, Options(seed=1, precision=MIXED))
    assert np.array_equal(lab64, lab32)
    assert eq_bits(rep64.S, rep32.S) and rep64.B == rep32.B and rep64.sweeps == rep32.sweeps
    assert rep32.fp64_evals > 0 and 0 < rep32.max_abs_err_fp32 < 1e-6 * abs(rep32.S)
    print(f"C3 fp64 {rep64.t_total:.3f}s mixed {rep32.t_total:.3f}s rechecks {rep32.rechecks} "
          f"fp64_evals {rep32.fp64_evals} err {rep32.max_abs_err_fp32:.3g}", flush=True)
