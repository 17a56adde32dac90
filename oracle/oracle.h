/* CPU oracle for the GSAP stochastic-block-partitioning hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; it is linked or loaded only by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The
 * product (gsap_amd/) never includes, links or calls anything here.
 *
 * It restates, single-threaded and in plain C++17, the algorithm of
 *   GSAP/partition.hpp   (control flow, merge order, golden section, convergence)
 *   GSAP/kernels.hpp     (every proposal / delta-MDL / MDL formula)
 * with the documented decisions on the reference's quirks (see oracle.cpp).
 *
 * Every function is exported with C linkage so that Python tests can drive it
 * through ctypes with plain pointers.
 */
#ifndef GSAP_ORACLE_H
#define GSAP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Canonical FP64 log2 (table of 128 + degree-9 polynomial, fma only). */
double orc_log2(double x);
/* The oracle's own copy of the canonical log2 constants (computed in binary128
 * at start-up, rounded once): inv_c[128], log2_c[128], poly[9]. */
void orc_log2_tables(double *inv_c, double *log2_c, double *poly);

/* Canonical 64-way sum of n terms (64 strided partials, then a binary tree). */
double orc_csum64(const double *x, int64_t n);
/* Canonical array sum: tiles of 1024 summed with csum64, tile sums with csum64. */
double orc_csum_array(const double *x, int64_t n);

/* Block matrix of a labelling.  Outputs (caller allocated):
 *   rowptr[B+1], col[E], val[E]   CSR, columns ascending, no zero entries
 *   colptr[B+1], row[E], cval[E]  CSC, rows ascending
 *   dout[B], din[B]               row / column sums
 * Returns nnz, or -1 on invalid input. */
int64_t orc_blockmodel(int64_t N, int64_t E, const int32_t *out_ptr, const int32_t *out_adj,
                       const int32_t *out_w, const int32_t *labels, int32_t B,
                       int64_t *rowptr, int32_t *col, int64_t *val,
                       int64_t *colptr, int32_t *row, int64_t *cval,
                       int64_t *dout, int64_t *din);

typedef struct orc_graph {
  int64_t N, E;
  const int32_t *out_ptr; /* N+1 */
  const int32_t *out_adj; /* E   */
  const int32_t *out_w;   /* E   */
  const int32_t *in_ptr;  /* N+1 */
  const int32_t *in_adj;  /* E   */
  const int32_t *in_w;    /* E   */
} orc_graph;

typedef struct orc_options {
  uint64_t seed;
  int32_t num_proposals;     /* 10   partition.hpp:855 */
  float reduction_rate;      /* 0.35 partition.hpp:856 */
  int32_t max_nodal_itr;     /* 100  partition.hpp:857 */
  int32_t num_batches;       /* 4    partition.hpp:858 */
  double threshold1;         /* 5e-4 partition.hpp:859 */
  double threshold2;         /* 1e-4 partition.hpp:860 */
  int32_t moving_avg_window; /* 3    partition.hpp:861 */
  int32_t max_levels;        /* safety cap on outer iterations */
  uint32_t chain;            /* Philox stream id (counter word 2) */
  int32_t model;             /* 0: model term as partition.hpp:696-699, 1: x ln x scaled by E */
  int32_t delta_mode;        /* 0: delta-S as kernels.hpp computes it, 1: exact change of S */
  int32_t accept;            /* 0: greedy dS < 0 (kernels.hpp:600-643), 1: Metropolis-Hastings */
  double beta;               /* inverse temperature of accept = 1 */
} orc_options;

typedef struct orc_trace_rec {
  int32_t kind; /* 1 merge, 2 batch, 3 itr, 4 golden, 5 rebuild */
  int32_t B;
  int32_t a, b;
  uint64_t h1, h2;
  double x, y;
} orc_trace_rec;

typedef struct orc_report {
  int32_t B;
  int32_t levels;
  int32_t merge_phases;
  int32_t sweeps;
  int64_t batches;
  double S;
  int64_t n_trace;
} orc_report;

/* Single merge phase on a given labelling.  Writes prop[np*B], dS[np*B],
 * best[B], dS_best[B], new labels (in place) and returns the new B.  m is the
 * merge-phase counter used in the Philox counter. */
int32_t orc_merge_phase(const orc_graph *g, const orc_options *o, int32_t *labels, int32_t B,
                        int32_t B_to_merge, uint32_t m, int32_t *prop, double *dS,
                        int32_t *best, double *dS_best);

/* Single nodal batch b (0..num_batches-1) at batch counter q.  Writes
 * prop[BS] (greedy: the proposals; Metropolis-Hastings: the label each vertex
 * ends with), dS[BS] (0 for rejected moves), updates labels in place and
 * returns the canonical sum of accepted delta-S in *dE. */
void orc_nodal_batch(const orc_graph *g, const orc_options *o, int32_t *labels, int32_t B,
                     int32_t b, uint32_t q, int32_t *prop, double *dS, double *dE);
/* The same batch with the per-vertex proposals / delta-S and the per-row sorts of the
 * block-matrix build spread over `threads` host threads (identical results). */
void orc_nodal_batch_mt(const orc_graph *g, const orc_options *o, int32_t *labels, int32_t B,
                        int32_t b, uint32_t q, int32_t *prop, double *dS, double *dE,
                        int32_t threads);

/* delta-S of moving vertex v from its block to s against the labelling's block matrix
 * (exact = 0: as kernels.hpp computes it; 1: the change of S).  With log2H, also the log2
 * Hastings ratio log2 p(s -> r) - log2 p(r -> s) of the nodal sampler. */
double orc_nodal_delta(const orc_graph *g, const int32_t *labels, int32_t B, int64_t v,
                       int32_t s, int32_t exact, double *log2H);
/* delta-S of merging block r into block s (exact as above). */
double orc_merge_delta(const orc_graph *g, const int32_t *labels, int32_t B, int32_t r,
                       int32_t s, int32_t exact);

/* Total description length of a labelling. */
double orc_mdl(const orc_graph *g, const int32_t *labels, int32_t B, int32_t model);

/* Whole partition().  labels_out[N]; trace may be NULL (cap = max records). */
int orc_partition(const orc_graph *g, const orc_options *o, int32_t *labels_out,
                  orc_report *rep, orc_trace_rec *trace, int64_t trace_cap);

/* One slot of the golden-section bracket: Partition (graph.hpp:48-54) with S in FP64 and
 * the id (0..2) of the label snapshot it owns. */
typedef struct orc_golden_slot {
  int32_t B;
  int32_t B_to_merge;
  double S;
  int32_t snap;
  int32_t pad;
} orc_golden_slot;

/* Slots large, med, small as OldData() leaves them (B = 0, S = inf); slot k owns snapshot k. */
void orc_golden_init(orc_golden_slot slots[3]);
/* Convergence test after sweep itr (partition.hpp:918-934): 1 when the level stops.
 * dE[0..itr] are the sweeps' accepted delta-S sums; bracketed = all three slots finite. */
int32_t orc_converged(const double *dE, int32_t itr, int32_t W, double S, int32_t bracketed,
                      double threshold1, double threshold2);
/* Golden-section bracket update after a level (prepare_for_partition_next,
 * partition.hpp:80-322).  slots = {large, med, small}, updated in place; P = the level's
 * result (B, B_to_merge, S).  *store = snapshot that receives the level's labels;
 * *restore = snapshot the next level starts from; *next = its B, next B_to_merge and S;
 * *optimal = 1 when the search ends. */
void orc_golden_step(orc_golden_slot slots[3], const orc_golden_slot *P, float rate,
                     int32_t *store, int32_t *restore, orc_golden_slot *next, int32_t *optimal);

/* The same partition with the nodal batches and block-matrix builds spread over `threads`
 * host threads (identical results). */
int orc_partition_mt(const orc_graph *g, const orc_options *o, int32_t *labels_out,
                     orc_report *rep, orc_trace_rec *trace, int64_t trace_cap, int32_t threads);

/* Greedy merge order + relabel (partition.hpp:26-78) as union-find; map_out[B] gets the
 * new label of every old block.  Returns the number of merges carried out. */
int32_t orc_carry_out_merges(const int32_t *best, const double *dS_best, int32_t B,
                             int32_t B_to_merge, int32_t *map_out);
/* find_best_warp's argmin (kernels.hpp:452-491) over np <= 32 proposals. */
void orc_argmin_tree(const double *dS, const int32_t *prop, int32_t np, double *best_dS,
                     int32_t *best);

/* Pairwise precision / recall of a labelling against the truth
 * (evaluate.hpp:130-181 definitions, int64 counts). */
void orc_pairwise(int64_t N, const int32_t *truth, const int32_t *given, double *precision,
                  double *recall);

/* FNV-1a 64 over a byte range (trace digests). */
uint64_t orc_fnv1a(const void *p, int64_t nbytes, uint64_t h);

#ifdef __cplusplus
}
#endif
#endif
