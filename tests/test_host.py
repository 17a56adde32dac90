"""CPU tests of the host-side pieces: TSV reader, DC-SBM generator, library exports."""
import ctypes
import os
import re

import numpy as np
import pytest

import gsap_amd
from conftest import GOLDEN, ROOT, tsv_graph
import oracle_lib as O


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "gsap", "gsap.h")).read()
    names = set(re.findall(r"\b(gsap_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 15
    L = gsap_amd.lib()
    for n in sorted(names):
        assert hasattr(L, n), n
    assert names == set(gsap_amd.EXPORTS)


def test_oracle_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "oracle", "oracle.h")).read()
    L = O.lib()
    for n in set(re.findall(r"\b(orc_[a-z0-9_]+)\s*\(", hdr)):
        assert hasattr(L, n), n


def test_tsv_reader_matches_file():
    path = os.path.join(GOLDEN, "graphchallenge", "LL", "1000_nodes.tsv")
    rows = [l.split() for l in open(path) if l.strip()]
    g = tsv_graph("LL")
    assert g.N == 1000 and g.E == len(rows)
    # the first line is "1 408<TAB>1": space-separated, read as 0 -> 407 (the reference's
    # tab-only parser reads it as a self-loop 1 -> 1, load.hpp:65-75)
    assert g.out_adj[g.out_ptr[0]] == 407
    src = np.array([int(r[0]) - 1 for r in rows]); dst = np.array([int(r[1]) - 1 for r in rows])
    for v in (0, 1, 500, 999):
        assert list(g.out_adj[g.out_ptr[v]:g.out_ptr[v + 1]]) == list(dst[src == v])
        assert list(g.in_adj[g.in_ptr[v]:g.in_ptr[v + 1]]) == list(src[dst == v])
    assert g.truth.min() == 0 and len(np.unique(g.truth)) == 11


def test_reader_rejects_missing_file():
    with pytest.raises(gsap_amd.GsapError):
        gsap_amd.read_tsv("/nonexistent/graph.tsv")


@pytest.mark.parametrize("alpha,rho,K", [(0.0, 4.0, 11), (1.0, 1.5, 33), (0.5, 2.0, 20)])
def test_generator_properties(alpha, rho, K):
    N = 4000
    g = gsap_amd.generate_dcsbm(N, K, alpha, rho, 20.0, seed=7)
    g2 = gsap_amd.generate_dcsbm(N, K, alpha, rho, 20.0, seed=7)
    for a, b in zip((g.out_ptr, g.out_adj, g.in_adj, g.truth), (g2.out_ptr, g2.out_adj, g2.in_adj, g2.truth)):
        assert np.array_equal(a, b)  # deterministic
    assert g.E == N * 20
    src = np.repeat(np.arange(N), np.diff(g.out_ptr))
    assert not np.any(src == g.out_adj)  # no self-loops
    keys = src.astype(np.int64) * N + g.out_adj
    assert len(np.unique(keys)) == g.E  # no duplicates
    # in-CSR is the transpose of out-CSR
    isrc = np.repeat(np.arange(N), np.diff(g.in_ptr))
    assert np.array_equal(np.sort(keys), np.sort(g.in_adj.astype(np.int64) * N + isrc))
    sizes = np.bincount(g.truth, minlength=K)
    assert sizes.min() >= 1 and sizes.sum() == N
    if alpha == 0:
        assert sizes.max() - sizes.min() <= 1
    intra = np.mean(g.truth[src] == g.truth[g.out_adj])
    # intra-block share rho/(1+rho) before rejections; allow sampling noise + rejection bias
    assert abs(intra - rho / (1 + rho)) < 0.03


def test_generator_pinned_digest():
    g = gsap_amd.generate_dcsbm(1000, 11, 0.0, 4.0, 20.0, seed=1)
    h = O.lib().orc_fnv1a
    h.restype = ctypes.c_uint64
    h.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64]
    d = h(g.out_adj.ctypes.data, 4 * g.E, 1469598103934665603)
    d = h(g.truth.ctypes.data, 4 * g.N, d)
    with open(os.path.join(GOLDEN, "dcsbm_c1_digest.txt")) as f:
        assert int(f.read().split()[0], 16) == d


def test_fp32_log2_error_bound():
    """The mixed form's FP32 log2 stays within (2 |log2 x| + 6) 2^-24 of the exact log2 for
    every integer count up to 2^24 and a sample up to 2^31 (the bound its delta-S error
    bound is built on)."""
    u = 2.0 ** -24
    x = np.arange(1, 2 ** 24 + 1, dtype=np.float64)
    L = np.log2(x)
    y = gsap_amd.log2f(x.astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(y - L) <= (2 * np.abs(L) + 6) * u)
    xi = np.random.default_rng(0).integers(2 ** 24, 2 ** 31, 1_000_000).astype(np.float64)
    L = np.log2(xi)
    y = gsap_amd.log2f(xi.astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(y - L) <= (2 * np.abs(L) + 6) * u)
    # powers of two are exact
    p = 2.0 ** np.arange(0, 31)
    assert np.array_equal(gsap_amd.log2f(p.astype(np.float32)), np.arange(0, 31, dtype=np.float32))


def test_csr_cache_round_trip(tmp_path):
    """Binary CSR cache: a generated graph and a TSV graph with truth come back identical;
    a flipped byte, a truncated file, a foreign file and a missing file are errors."""
    g = gsap_amd.generate_dcsbm(3000, 9, 1.0, 2.0, 15.0, seed=11)
    p = str(tmp_path / "g.csr")
    gsap_amd.save_csr(g, p)
    h = gsap_amd.load_csr(p)
    assert h.N == g.N and h.E == g.E
    for a in ("out_ptr", "out_adj", "out_w", "in_ptr", "in_adj", "in_w", "truth"):
        assert np.array_equal(getattr(h, a), getattr(g, a)), a
    t = tsv_graph("LH")
    q = str(tmp_path / "t.csr")
    gsap_amd.save_csr(t, q)
    u = gsap_amd.load_csr(q)
    assert np.array_equal(u.truth, t.truth) and np.array_equal(u.in_adj, t.in_adj)
    # graph without truth
    t.truth = None
    gsap_amd.save_csr(t, q)
    assert gsap_amd.load_csr(q).truth is None
    raw = bytearray(open(p, "rb").read())
    bad = str(tmp_path / "bad.csr")
    raw2 = bytearray(raw); raw2[len(raw2) // 2] ^= 0x40
    open(bad, "wb").write(raw2)
    with pytest.raises(gsap_amd.GsapError):
        gsap_amd.load_csr(bad)
    open(bad, "wb").write(raw[:len(raw) - 100])
    with pytest.raises(gsap_amd.GsapError):
        gsap_amd.load_csr(bad)
    open(bad, "wb").write(b"not a graph file at all, just text" * 4)
    with pytest.raises(gsap_amd.GsapError):
        gsap_amd.load_csr(bad)
    with pytest.raises(gsap_amd.GsapError):
        gsap_amd.load_csr(str(tmp_path / "missing.csr"))
