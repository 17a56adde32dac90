"""gsap_amd -- MI355X-native stochastic block partitioning (the GSAP hot path).

Thin ctypes binding of the C ABI declared in include/gsap/gsap.h and implemented by the
in-tree libgsap_amd.so (built by gsap_amd/build.py).  There is no fallback: if the
library is missing or a device call fails, a GsapError is raised.

Reference interface mirrored: GSAP/partition.hpp:852 `partition(Graph&)`; the Graph here
holds the same two CSR views (GSAP/graph.hpp:16-45) plus an explicit end pointer.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

__all__ = [
    "GsapError", "Graph", "Options", "Report", "Context", "lib", "generate_dcsbm", "read_tsv",
    "pairwise", "log2", "log2_tables", "PRECISION_FP64", "PRECISION_MIXED", "DELTA_REFERENCE",
    "DELTA_EXACT", "ACCEPT_GREEDY", "ACCEPT_MH", "GoldenSection", "converged",
]

PRECISION_FP64 = 0
PRECISION_MIXED = 1
DELTA_REFERENCE = 0
DELTA_EXACT = 1
ACCEPT_GREEDY = 0
ACCEPT_MH = 1
_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsap_amd.so")


class GsapError(RuntimeError):
    pass


_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class CGraph(C.Structure):
    _fields_ = [("N", C.c_int64), ("E", C.c_int64),
                ("out_ptr", C.c_void_p), ("out_adj", C.c_void_p), ("out_w", C.c_void_p),
                ("in_ptr", C.c_void_p), ("in_adj", C.c_void_p), ("in_w", C.c_void_p)]


class COptions(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("num_proposals", C.c_int32),
                ("reduction_rate", C.c_float), ("max_nodal_itr", C.c_int32),
                ("num_batches", C.c_int32), ("threshold1", C.c_double),
                ("threshold2", C.c_double), ("moving_avg_window", C.c_int32),
                ("max_levels", C.c_int32), ("precision", C.c_int32), ("chain", C.c_uint32),
                ("trace", C.c_int32), ("dense_max_B", C.c_int32), ("profile", C.c_int32),
                ("model", C.c_int32), ("workers", C.c_int32),
                ("profile_classes", C.c_int32), ("graph", C.c_int32),
                ("tile_max_B", C.c_int32), ("delta_mode", C.c_int32), ("accept", C.c_int32),
                ("beta", C.c_double), ("nodal_max16", C.c_int32), ("nodal_max32", C.c_int32),
                ("nodal_maxwave", C.c_int32), ("seg_reg_max", C.c_int32),
                ("seg_count_min", C.c_int32), ("sparse_update", C.c_int32),
                ("csr_every_batch", C.c_int32)]


class CReport(C.Structure):
    _fields_ = [("B", C.c_int32), ("levels", C.c_int32), ("merge_phases", C.c_int32),
                ("sweeps", C.c_int32), ("batches", C.c_int64), ("S", C.c_double),
                ("t_total", C.c_double), ("t_merge", C.c_double), ("t_sweep", C.c_double),
                ("t_rebuild", C.c_double), ("sweep_edges", C.c_int64),
                ("max_abs_err_fp32", C.c_double), ("rechecks", C.c_int64),
                ("n_trace", C.c_int64), ("kernel_ms", C.c_double * 8),
                ("kernel_calls", C.c_int64 * 8), ("fp64_evals", C.c_int64),
                ("t_upload", C.c_double), ("nodal_evals", C.c_int64),
                ("nodal_eval_adj", C.c_int64)]


class CQuality(C.Structure):
    _fields_ = [("n", C.c_int64), ("B_truth", C.c_int32), ("B_given", C.c_int32),
                ("accuracy", C.c_double), ("precision", C.c_double), ("recall", C.c_double),
                ("rand_index", C.c_double), ("adjusted_rand_index", C.c_double),
                ("nmi", C.c_double), ("accuracy_computed", C.c_int32), ("reserved", C.c_int32)]


class CTrace(C.Structure):
    _fields_ = [("kind", C.c_int32), ("B", C.c_int32), ("a", C.c_int32), ("b", C.c_int32),
                ("h1", C.c_uint64), ("h2", C.c_uint64), ("x", C.c_double), ("y", C.c_double)]


class CGoldenSlot(C.Structure):
    _fields_ = [("B", C.c_int32), ("B_to_merge", C.c_int32), ("S", C.c_double),
                ("snap", C.c_int32), ("reserved", C.c_int32)]


class CDcsbm(C.Structure):
    _fields_ = [("N", C.c_int64), ("K", C.c_int32), ("alpha", C.c_double), ("rho", C.c_double),
                ("mean_deg", C.c_double), ("deg_min", C.c_int32), ("deg_max", C.c_int32),
                ("exponent", C.c_double), ("seed", C.c_uint64)]


EXPORTS = [
    "gsap_options_default", "gsap_create", "gsap_destroy", "gsap_last_error",
    "gsap_partition_i32", "gsap_get_trace", "gsap_blockmodel_i32", "gsap_merge_phase_i32",
    "gsap_nodal_batch_i32", "gsap_nodal_batch_bm_i32", "gsap_mdl_i32", "gsap_generate_dcsbm_i32", "gsap_read_tsv_i32",
    "gsap_pairwise_i32", "gsap_log2", "gsap_log2_tables", "gsap_partition_batch_i32",
    "gsap_partition_chains_i32", "gsap_log2f_n", "gsap_quality_i32", "gsap_save_csr_i32",
    "gsap_load_csr_i32", "gsap_autotune_i32", "gsap_golden_init", "gsap_golden_step",
    "gsap_converged", "gsap_mdl_noalloc_i32", "gsap_upload_graph_i32",
    "gsap_partition_resident_i32",
]

# the option fields the autotuner writes
TUNED = ("tile_max_B", "dense_max_B", "nodal_max16", "nodal_max32", "nodal_maxwave",
         "seg_reg_max", "seg_count_min")

_lib = None


def lib() -> C.CDLL:
    """Loads libgsap_amd.so (raises GsapError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GsapError(f"{LIB_PATH} not found: run `python -m gsap_amd.build` first")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, u32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32
    L.gsap_options_default.argtypes = [C.POINTER(COptions)]
    L.gsap_options_default.restype = None
    L.gsap_create.argtypes = [C.POINTER(vp), C.c_int]
    L.gsap_destroy.argtypes = [vp]
    L.gsap_last_error.argtypes = [vp]
    L.gsap_last_error.restype = C.c_char_p
    L.gsap_partition_i32.argtypes = [vp, C.POINTER(CGraph), C.POINTER(COptions), _i32p,
                                     C.POINTER(CReport)]
    L.gsap_upload_graph_i32.argtypes = [vp, C.POINTER(CGraph)]
    L.gsap_partition_resident_i32.argtypes = [vp, C.POINTER(COptions), C.c_int64, _i32p,
                                              C.POINTER(CReport)]
    L.gsap_get_trace.argtypes = [vp, C.POINTER(CTrace), i64, C.POINTER(i64)]
    L.gsap_partition_batch_i32.argtypes = [vp, C.POINTER(CGraph), i32, C.POINTER(COptions),
                                           C.POINTER(C.c_void_p), C.POINTER(CReport)]
    L.gsap_partition_chains_i32.argtypes = [vp, C.POINTER(CGraph), C.POINTER(COptions), i32,
                                            _i32p, C.POINTER(CReport), _f64p, C.POINTER(i32)]
    L.gsap_blockmodel_i32.argtypes = [vp, C.POINTER(CGraph), _i32p, i32, _i64p, _i32p, _i64p,
                                      _i64p, _i32p, _i64p, _i64p, _i64p, C.POINTER(i64),
                                      C.POINTER(COptions)]
    L.gsap_nodal_batch_bm_i32.argtypes = [vp, C.POINTER(CGraph), C.POINTER(COptions), _i32p,
                                          i32, i32, u32, i32, _i64p, _i32p, _i64p, _i64p,
                                          _i32p, _i64p, _i64p, _i64p, C.POINTER(i64)]
    L.gsap_merge_phase_i32.argtypes = [vp, C.POINTER(CGraph), C.POINTER(COptions), _i32p, i32,
                                       i32, u32, _i32p, _f64p, _i32p, _f64p, C.POINTER(i32)]
    L.gsap_nodal_batch_i32.argtypes = [vp, C.POINTER(CGraph), C.POINTER(COptions), _i32p, i32,
                                       i32, u32, _i32p, _f64p, C.POINTER(C.c_double)]
    L.gsap_mdl_i32.argtypes = [vp, C.POINTER(CGraph), _i32p, i32, C.POINTER(C.c_double),
                               C.POINTER(COptions)]
    L.gsap_generate_dcsbm_i32.argtypes = [C.POINTER(CDcsbm), C.POINTER(i64), vp, vp, vp, vp, vp,
                                          vp, vp]
    L.gsap_read_tsv_i32.argtypes = [C.c_char_p, C.POINTER(i64), C.POINTER(i64), vp, vp, vp, vp,
                                    vp, vp]
    L.gsap_pairwise_i32.argtypes = [i64, _i32p, _i32p, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
    L.gsap_quality_i32.argtypes = [i64, _i32p, _i32p, C.POINTER(CQuality)]
    L.gsap_autotune_i32.argtypes = [vp, C.POINTER(CGraph), C.POINTER(COptions)]
    L.gsap_save_csr_i32.argtypes = [C.c_char_p, C.POINTER(CGraph), vp]
    L.gsap_load_csr_i32.argtypes = [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int32)] + [vp] * 7
    L.gsap_log2.argtypes = [C.c_double]
    L.gsap_log2.restype = C.c_double
    L.gsap_log2_tables.argtypes = [_f64p, _f64p, _f64p]
    L.gsap_log2_tables.restype = None
    L.gsap_log2f_n.argtypes = [_f32p, _f32p, C.c_int64]
    L.gsap_log2f_n.restype = None
    L.gsap_mdl_noalloc_i32.argtypes = [vp, vp, vp, vp, vp, i64, i64, i32, i32, vp, vp]
    L.gsap_golden_init.argtypes = [C.POINTER(CGoldenSlot)]
    L.gsap_golden_init.restype = None
    L.gsap_golden_step.argtypes = [C.POINTER(CGoldenSlot), C.POINTER(CGoldenSlot), C.c_float,
                                   C.POINTER(i32), C.POINTER(i32), C.POINTER(CGoldenSlot),
                                   C.POINTER(i32)]
    L.gsap_golden_step.restype = None
    L.gsap_converged.argtypes = [_f64p, i32, i32, C.c_double, i32, C.c_double, C.c_double]
    L.gsap_converged.restype = i32
    _lib = L
    return L


# ---------------------------------------------------------------------- data types
@dataclass
class Graph:
    """Directed weighted graph in out- and in-CSR form (0-based int32 ids, N+1 ptrs)."""
    N: int
    out_ptr: np.ndarray
    out_adj: np.ndarray
    out_w: np.ndarray
    in_ptr: np.ndarray
    in_adj: np.ndarray
    in_w: np.ndarray
    truth: Optional[np.ndarray] = None

    @property
    def E(self) -> int:
        return int(self.out_adj.shape[0])

    def cstruct(self) -> CGraph:
        for a in (self.out_ptr, self.out_adj, self.out_w, self.in_ptr, self.in_adj, self.in_w):
            assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
        return CGraph(self.N, self.E, self.out_ptr.ctypes.data, self.out_adj.ctypes.data,
                      self.out_w.ctypes.data, self.in_ptr.ctypes.data, self.in_adj.ctypes.data,
                      self.in_w.ctypes.data)

    @staticmethod
    def from_edges(N: int, src, dst, w=None) -> "Graph":
        """CSR views of an edge list; rows keep the edge-list order (GSAP/load.hpp:80-111)."""
        src = np.asarray(src, dtype=np.int64)
        dst = np.asarray(dst, dtype=np.int64)
        w = np.ones(len(src), dtype=np.int32) if w is None else np.asarray(w, dtype=np.int32)

        def csr(a, b):
            order = np.argsort(a, kind="stable")
            ptr = np.zeros(N + 1, dtype=np.int32)
            np.add.at(ptr, a + 1, 1)
            ptr = np.cumsum(ptr).astype(np.int32)
            return ptr, b[order].astype(np.int32), w[order].astype(np.int32)

        op, oa, ow = csr(src, dst)
        ip, ia, iw = csr(dst, src)
        return Graph(N, op, oa, ow, ip, ia, iw)


@dataclass
class Options:
    """Algorithm options; defaults are GSAP/partition.hpp:855-861."""
    seed: int = 1
    num_proposals: int = 10
    reduction_rate: float = 0.35
    max_nodal_itr: int = 100
    num_batches: int = 4
    threshold1: float = 5e-4
    threshold2: float = 1e-4
    moving_avg_window: int = 3
    max_levels: int = 0
    precision: int = PRECISION_FP64
    chain: int = 0
    trace: int = 0
    dense_max_B: int = 0
    profile: int = 0
    profile_classes: int = 0
    model: int = 0  # 0: the reference's model term, 1: corrected (x ln x scaled by E)
    workers: int = 0  # batch / chain calls: concurrent sub-contexts (0 = 8)
    graph: int = 0  # 1: each level's sweep captured once and replayed as a HIP graph
    tile_max_B: int = 0  # row-tile rebuild up to this B (0 = auto)
    delta_mode: int = 0  # DELTA_REFERENCE or DELTA_EXACT
    accept: int = 0  # ACCEPT_GREEDY or ACCEPT_MH
    beta: float = 3.0  # inverse temperature of ACCEPT_MH
    # work-split choices (identical results; 0 = default): nodal neighbour-count classes and
    # segment-sort classes of the segmented rebuild
    nodal_max16: int = 0
    nodal_max32: int = 0
    nodal_maxwave: int = 0
    seg_reg_max: int = 0
    seg_count_min: int = 0
    # after a batch without the dense block matrix: 0 rebuilds from every edge, 1 merges the
    # moves into the current CSR/CSC (identical results)
    sparse_update: int = 0
    # at a dense B: 0 rebuilds the CSR/CSC after a sweep's last batch only (the batches
    # before it update the dense matrix and the block degrees), 1 after every batch
    # (identical results)
    csr_every_batch: int = 0

    def cstruct(self) -> COptions:
        o = COptions()
        for f in ("seed", "num_proposals", "reduction_rate", "max_nodal_itr", "num_batches",
                  "threshold1", "threshold2", "moving_avg_window", "max_levels", "precision",
                  "chain", "trace", "dense_max_B", "profile", "model", "workers",
                  "profile_classes", "graph", "tile_max_B", "delta_mode", "accept", "beta",
                  "nodal_max16", "nodal_max32", "nodal_maxwave", "seg_reg_max",
                  "seg_count_min", "sparse_update", "csr_every_batch"):
            setattr(o, f, getattr(self, f))
        return o


@dataclass
class Report:
    B: int
    levels: int
    merge_phases: int
    sweeps: int
    batches: int
    S: float
    t_total: float
    t_merge: float
    t_sweep: float
    sweep_edges: int
    max_abs_err_fp32: float
    rechecks: int
    n_trace: int
    kernel_ms: list = field(default_factory=list)
    kernel_calls: list = field(default_factory=list)
    fp64_evals: int = 0
    t_upload: float = 0.0
    nodal_evals: int = -1
    nodal_eval_adj: int = -1

    @staticmethod
    def from_c(r: CReport) -> "Report":
        return Report(r.B, r.levels, r.merge_phases, r.sweeps, r.batches, r.S, r.t_total,
                      r.t_merge, r.t_sweep, r.sweep_edges, r.max_abs_err_fp32, r.rechecks,
                      r.n_trace, list(r.kernel_ms), list(r.kernel_calls), r.fp64_evals,
                      r.t_upload, r.nodal_evals, r.nodal_eval_adj)


def _check(rc: int, ctx=None):
    if rc != 0:
        msg = lib().gsap_last_error(ctx).decode() if ctx else ""
        raise GsapError(f"gsap error {rc}: {msg}")


class Context:
    """Owns the device memory of one GPU; one per host thread."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        rc = lib().gsap_create(C.byref(self._h), device)
        if rc != 0:
            raise GsapError(f"gsap_create(device={device}) failed with {rc}")

    def close(self):
        if self._h:
            lib().gsap_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def partition(self, g: Graph, opts: Options = Options()):
        out = np.empty(g.N, dtype=np.int32)
        rep = CReport()
        cg, co = g.cstruct(), opts.cstruct()
        _check(lib().gsap_partition_i32(self._h, C.byref(cg), C.byref(co), out, C.byref(rep)),
               self._h)
        return out, Report.from_c(rep)

    def upload_graph(self, g: Graph) -> None:
        """Copies g into the context's device memory (the first step of partition())."""
        cg = g.cstruct()
        _check(lib().gsap_upload_graph_i32(self._h, C.byref(cg)), self._h)

    def partition_resident(self, N: int, opts: Options = Options()):
        """Whole partition of the N-vertex graph resident on this context (left by
        upload_graph() or by any call that took a graph), with no host-to-device copy of
        it; the same result as partition() on that graph."""
        out = np.empty(N, dtype=np.int32)
        rep = CReport()
        co = opts.cstruct()
        _check(lib().gsap_partition_resident_i32(self._h, C.byref(co), N, out, C.byref(rep)),
               self._h)
        return out, Report.from_c(rep)

    def autotune(self, g: Graph, opts: Options = Options()) -> Options:
        """Options with the path thresholds and work-split classes measured on this graph
        and device (identical results, other fields kept)."""
        import dataclasses
        cg, co = g.cstruct(), opts.cstruct()
        _check(lib().gsap_autotune_i32(self._h, C.byref(cg), C.byref(co)), self._h)
        return dataclasses.replace(opts, **{f: getattr(co, f) for f in TUNED})

    def partition_batch(self, graphs, opts: Options = Options()):
        """Batched form: independent graphs on concurrent sub-contexts."""
        n = len(graphs)
        cg = (CGraph * max(n, 1))(*[g.cstruct() for g in graphs])
        outs = [np.empty(g.N, dtype=np.int32) for g in graphs]
        ptrs = (C.c_void_p * max(n, 1))(*[o.ctypes.data for o in outs])
        reps = (CReport * max(n, 1))()
        co = opts.cstruct()
        _check(lib().gsap_partition_batch_i32(self._h, cg, n, C.byref(co), ptrs, reps), self._h)
        return outs, [Report.from_c(reps[k]) for k in range(n)]

    def partition_chains(self, g: Graph, n_chains: int, opts: Options = Options()):
        """n_chains chains (stream ids opts.chain + k); returns the lowest-S labelling."""
        out = np.empty(g.N, dtype=np.int32)
        rep = CReport()
        S = np.zeros(n_chains, np.float64)
        best = C.c_int32()
        cg, co = g.cstruct(), opts.cstruct()
        _check(lib().gsap_partition_chains_i32(self._h, C.byref(cg), C.byref(co), n_chains, out,
                                               C.byref(rep), S, C.byref(best)), self._h)
        return out, Report.from_c(rep), S, best.value

    def trace(self):
        n = C.c_int64()
        _check(lib().gsap_get_trace(self._h, None, 0, C.byref(n)), self._h)
        buf = (CTrace * max(n.value, 1))()
        _check(lib().gsap_get_trace(self._h, buf, n.value, C.byref(n)), self._h)
        return [(t.kind, t.B, t.a, t.b, t.h1, t.h2, t.x, t.y) for t in buf[:n.value]]

    def blockmodel(self, g: Graph, labels: np.ndarray, B: int, opts: Optional[Options] = None):
        E = g.E
        rowptr = np.zeros(B + 1, np.int64); colptr = np.zeros(B + 1, np.int64)
        col = np.zeros(max(E, 1), np.int32); row = np.zeros(max(E, 1), np.int32)
        val = np.zeros(max(E, 1), np.int64); cval = np.zeros(max(E, 1), np.int64)
        dout = np.zeros(B, np.int64); din = np.zeros(B, np.int64)
        nnz = C.c_int64()
        cg = g.cstruct()
        co = (opts or Options()).cstruct()
        lab = np.ascontiguousarray(labels, dtype=np.int32)
        _check(lib().gsap_blockmodel_i32(self._h, C.byref(cg), lab, B, rowptr, col, val, colptr,
                                         row, cval, dout, din, C.byref(nnz), C.byref(co)),
               self._h)
        k = nnz.value
        return dict(rowptr=rowptr, col=col[:k], val=val[:k], colptr=colptr, row=row[:k],
                    cval=cval[:k], dout=dout, din=din, nnz=k)

    def merge_phase(self, g: Graph, opts: Options, labels: np.ndarray, B: int, B_to_merge: int,
                    m: int):
        np_ = opts.num_proposals
        lab = np.ascontiguousarray(labels, dtype=np.int32).copy()
        prop = np.zeros(np_ * B, np.int32); dS = np.zeros(np_ * B, np.float64)
        best = np.zeros(B, np.int32); dSb = np.zeros(B, np.float64)
        bn = C.c_int32()
        cg, co = g.cstruct(), opts.cstruct()
        _check(lib().gsap_merge_phase_i32(self._h, C.byref(cg), C.byref(co), lab, B, B_to_merge,
                                          m, prop, dS, best, dSb, C.byref(bn)), self._h)
        return dict(labels=lab, prop=prop, dS=dS, best=best, dS_best=dSb, B=bn.value)

    def nodal_batch(self, g: Graph, opts: Options, labels: np.ndarray, B: int, b: int, q: int):
        BS = g.N // opts.num_batches
        lab = np.ascontiguousarray(labels, dtype=np.int32).copy()
        prop = np.zeros(max(BS, 1), np.int32); dS = np.zeros(max(BS, 1), np.float64)
        dE = C.c_double()
        cg, co = g.cstruct(), opts.cstruct()
        _check(lib().gsap_nodal_batch_i32(self._h, C.byref(cg), C.byref(co), lab, B, b, q, prop,
                                          dS, C.byref(dE)), self._h)
        return dict(labels=lab, prop=prop[:BS], dS=dS[:BS], dE=dE.value)

    def nodal_batch_bm(self, g: Graph, opts: Options, labels: np.ndarray, B: int, b: int,
                       q: int, keep: bool = False):
        """Nodal batch b, then the block-matrix update the partition loop runs after it; the
        labels and block matrix afterwards.  keep: continue from the labels and block matrix
        the previous call left on this context (labels is then only the output's shape)."""
        E = g.E
        lab = np.ascontiguousarray(labels, dtype=np.int32).copy()
        rowptr = np.zeros(B + 1, np.int64); colptr = np.zeros(B + 1, np.int64)
        col = np.zeros(max(E, 1), np.int32); row = np.zeros(max(E, 1), np.int32)
        val = np.zeros(max(E, 1), np.int64); cval = np.zeros(max(E, 1), np.int64)
        dout = np.zeros(B, np.int64); din = np.zeros(B, np.int64)
        nnz = C.c_int64()
        cg, co = g.cstruct(), opts.cstruct()
        _check(lib().gsap_nodal_batch_bm_i32(self._h, C.byref(cg), C.byref(co), lab, B, b, q,
                                             1 if keep else 0, rowptr, col, val, colptr, row,
                                             cval, dout, din, C.byref(nnz)), self._h)
        k = nnz.value
        return lab, dict(rowptr=rowptr, col=col[:k], val=val[:k], colptr=colptr, row=row[:k],
                         cval=cval[:k], dout=dout, din=din, nnz=k)

    def mdl_noalloc(self, d_out_ptr: int, d_out_adj: int, d_out_w: int, d_labels: int, N: int,
                    E: int, B: int, d_S: int, stream: int = 0, model: int = 0) -> None:
        """gsap_mdl_noalloc_i32 on device buffers given by address (no allocation; the
        result is written to d_S on `stream`, not waited for)."""
        _check(lib().gsap_mdl_noalloc_i32(self._h, d_out_ptr, d_out_adj, d_out_w, d_labels, N,
                                          E, B, model, d_S, stream or None), self._h)

    def mdl(self, g: Graph, labels: np.ndarray, B: int, opts: Optional[Options] = None) -> float:
        S = C.c_double()
        cg = g.cstruct()
        co = (opts or Options()).cstruct()
        lab = np.ascontiguousarray(labels, dtype=np.int32)
        _check(lib().gsap_mdl_i32(self._h, C.byref(cg), lab, B, C.byref(S), C.byref(co)), self._h)
        return S.value


# ---------------------------------------------------------------- host helpers
class GoldenSection:
    """The golden-section search over the block count as the library runs it
    (prepare_for_partition_next, GSAP/partition.hpp:80-322; OldData, GSAP/graph.hpp:57-70).
    slots() gives (B, B_to_merge, S, snapshot) of large, med and small."""

    def __init__(self, rate: float = 0.35):
        self.rate = rate
        self._s = (CGoldenSlot * 3)()
        lib().gsap_golden_init(self._s)

    def slots(self):
        return [(x.B, x.B_to_merge, x.S, x.snap) for x in self._s]

    def step(self, B: int, B_to_merge: int, S: float):
        """Records a level's result; returns (store, restore, (B, B_to_merge, S), optimal):
        the snapshot that takes the level's labels, the one the next level starts from,
        that level's state, and whether the search ends."""
        P = CGoldenSlot(B, B_to_merge, S, -1, 0)
        nxt = CGoldenSlot()
        st, rs, opt = C.c_int32(), C.c_int32(), C.c_int32()
        lib().gsap_golden_step(self._s, C.byref(P), self.rate, C.byref(st), C.byref(rs),
                               C.byref(nxt), C.byref(opt))
        return st.value, rs.value, (nxt.B, nxt.B_to_merge, nxt.S), bool(opt.value)


def converged(dE, itr: int, S: float, bracketed: bool, W: int = 3, threshold1: float = 5e-4,
              threshold2: float = 1e-4) -> bool:
    """The library's convergence test after sweep itr (GSAP/partition.hpp:918-934)."""
    d = np.ascontiguousarray(dE, dtype=np.float64)
    return bool(lib().gsap_converged(d, itr, W, S, int(bracketed), threshold1, threshold2))


def generate_dcsbm(N: int, K: int, alpha: float, rho: float, mean_deg: float, seed: int,
                   deg_min: int = 5, deg_max: int = 100, exponent: float = 2.5) -> Graph:
    """Seeded degree-corrected SBM (the benchmark input generator)."""
    p = CDcsbm(N, K, alpha, rho, mean_deg, deg_min, deg_max, exponent, seed)
    E = C.c_int64()
    _check(lib().gsap_generate_dcsbm_i32(C.byref(p), C.byref(E), None, None, None, None, None,
                                         None, None))
    E = E.value
    arrs = [np.empty(N + 1, np.int32), np.empty(E, np.int32), np.empty(E, np.int32),
            np.empty(N + 1, np.int32), np.empty(E, np.int32), np.empty(E, np.int32),
            np.empty(N, np.int32)]
    _check(lib().gsap_generate_dcsbm_i32(C.byref(p), C.byref(C.c_int64()),
                                         *[a.ctypes.data for a in arrs]))
    return Graph(N, *arrs[:6], truth=arrs[6])


def read_tsv(path: str, truth_path: Optional[str] = None) -> Graph:
    """GraphChallenge TSV reader (1-based `src dst w`, any whitespace)."""
    N, E = C.c_int64(), C.c_int64()
    _check(lib().gsap_read_tsv_i32(path.encode(), C.byref(N), C.byref(E), None, None, None, None,
                                   None, None))
    n, e = N.value, E.value
    arrs = [np.empty(n + 1, np.int32), np.empty(e, np.int32), np.empty(e, np.int32),
            np.empty(n + 1, np.int32), np.empty(e, np.int32), np.empty(e, np.int32)]
    _check(lib().gsap_read_tsv_i32(path.encode(), C.byref(N), C.byref(E),
                                   *[a.ctypes.data for a in arrs]))
    g = Graph(n, *arrs)
    if truth_path:
        t = np.loadtxt(truth_path, dtype=np.int64, ndmin=2)
        truth = np.full(n, -1, np.int32)
        truth[t[:, 0] - 1] = t[:, 1] - 1
        g.truth = truth
    return g


def save_csr(g: Graph, path: str) -> None:
    """Binary CSR cache of the graph (and its truth labelling, when it has one)."""
    cg = g.cstruct()
    t = None
    if g.truth is not None:
        t = np.ascontiguousarray(g.truth, np.int32)
    _check(lib().gsap_save_csr_i32(path.encode(), C.byref(cg),
                                   t.ctypes.data if t is not None else None))


def load_csr(path: str) -> Graph:
    """Graph written by save_csr."""
    N, E, ht = C.c_int64(), C.c_int64(), C.c_int32()
    _check(lib().gsap_load_csr_i32(path.encode(), C.byref(N), C.byref(E), C.byref(ht),
                                   *([None] * 7)))
    n, e = N.value, E.value
    arrs = [np.empty(n + 1, np.int32), np.empty(e, np.int32), np.empty(e, np.int32),
            np.empty(n + 1, np.int32), np.empty(e, np.int32), np.empty(e, np.int32)]
    truth = np.empty(n, np.int32) if ht.value else None
    _check(lib().gsap_load_csr_i32(path.encode(), C.byref(N), C.byref(E), C.byref(ht),
                                   *[a.ctypes.data for a in arrs],
                                   truth.ctypes.data if truth is not None else None))
    g = Graph(n, *arrs)
    g.truth = truth
    return g


def pairwise(truth: np.ndarray, given: np.ndarray):
    """(precision, recall) of pair co-membership, GSAP/evaluate/evaluate.hpp:130-181."""
    p, r = C.c_double(), C.c_double()
    t = np.ascontiguousarray(truth, np.int32)
    g = np.ascontiguousarray(given, np.int32)
    _check(lib().gsap_pairwise_i32(len(t), t, g, C.byref(p), C.byref(r)))
    return p.value, r.value


def quality(truth: np.ndarray, given: np.ndarray) -> dict:
    """Full quality report (GSAP/evaluate/evaluate.hpp:19-265): optimal-matching accuracy,
    pairwise precision / recall, Rand and adjusted Rand index, NMI; accuracy is None when
    the contingency table is too large to match (more than 64M cells)."""
    q = CQuality()
    t = np.ascontiguousarray(truth, np.int32)
    g = np.ascontiguousarray(given, np.int32)
    _check(lib().gsap_quality_i32(len(t), t, g, C.byref(q)))
    return {"n": q.n, "B_truth": q.B_truth, "B_given": q.B_given,
            "accuracy": q.accuracy if q.accuracy_computed else None,
            "precision": q.precision, "recall": q.recall, "rand_index": q.rand_index,
            "adjusted_rand_index": q.adjusted_rand_index, "nmi": q.nmi}


def log2(x: float) -> float:
    return lib().gsap_log2(float(x))


def log2f(x) -> np.ndarray:
    """FP32 log2 of the mixed form (x >= 1), evaluated on the host by the device code."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().gsap_log2f_n(x, y, len(x))
    return y


def log2_tables():
    inv = np.zeros(128); lg = np.zeros(128); poly = np.zeros(9)
    lib().gsap_log2_tables(inv, lg, poly)
    return inv, lg, poly
