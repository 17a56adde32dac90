/* gsap.h -- C ABI of the MI355X-native stochastic-block-partitioning hot path.
 *
 * Plain pointers and sizes only; no exceptions cross this boundary.  Every call
 * returns an int status (GSAP_OK == 0) and leaves a message readable through
 * gsap_last_error().  A context owns all device memory; it is not thread-safe
 * (use one context per host thread).
 *
 * Reference interfaces these entry points replace (paths relative to the
 * reference checkout):
 *   gsap_partition_i32      <- template<T,V,W> void partition(Graph<T,V,W>&)
 *                              GSAP/partition.hpp:852-954, instantiated as
 *                              <int,int,unsigned> by GSAP.cu:20,25
 *   gsap_blockmodel_i32     <- calculate_new_csr, GSAP/partition.hpp:707-849
 *   gsap_merge_phase_i32    <- propose_block_merge + carry_out_best_merges,
 *                              GSAP/partition.hpp:324-454 and 26-78
 *   gsap_nodal_batch_i32    <- propose_nodal_move, GSAP/partition.hpp:456-662
 *   gsap_nodal_batch_bm_i32 <- propose_nodal_move + calculate_new_csr,
 *                              GSAP/partition.hpp:914-915
 *   gsap_mdl_i32            <- compute_overall_entropy_cuda, GSAP/partition.hpp:665-704
 *   gsap_converged          <- the convergence test of GSAP/partition.hpp:918-934
 *   gsap_golden_step        <- prepare_for_partition_next, GSAP/partition.hpp:80-322
 *   gsap_golden_init        <- OldData(), GSAP/graph.hpp:57-70
 *   gsap_read_tsv_i32       <- readFile, GSAP/load.hpp:6-133 (whitespace tolerant)
 *   gsap_pairwise_i32       <- bf::evaluate pair counts, GSAP/evaluate/evaluate.hpp:130-181
 *   gsap_quality_i32        <- bf::evaluate (accuracy, pair counts, NMI),
 *                              GSAP/evaluate/evaluate.hpp:19-265, km.hpp:19-115
 *   gsap_generate_dcsbm_i32 -- no counterpart: the seeded DC-SBM generator that
 *                              produces the benchmark graphs.
 */
#ifndef GSAP_GSAP_H
#define GSAP_GSAP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSAP_OK 0
#define GSAP_ERR_INVALID 1   /* bad argument or malformed graph */
#define GSAP_ERR_DEVICE 2    /* HIP runtime error */
#define GSAP_ERR_MEMORY 3    /* device allocation failed */
#define GSAP_ERR_INTERNAL 4  /* invariant broken inside the library */
#define GSAP_ERR_IO 5        /* file could not be read */

#define GSAP_PRECISION_FP64 0  /* plain form: FP64 delta-S everywhere */
#define GSAP_PRECISION_MIXED 1 /* FP32 bulk delta-S + FP64 near-threshold recheck */

/* Model term of the description length (partition.hpp:696-699):
 *   REFERENCE  E(1+x)ln(1+x) - x ln x + N ln B,   x = B^2/E  (as the reference computes it)
 *   CORRECTED  E[(1+x)ln(1+x) - x ln x] + N ln B              (the x ln x term scaled by E) */
#define GSAP_MODEL_REFERENCE 0
#define GSAP_MODEL_CORRECTED 1

/* delta-S of nodal moves and block merges */
#define GSAP_DELTA_REFERENCE 0 /* as the reference computes it (kernels.hpp:233-320,
                                  1444-1726: new rows keep the old column degrees of r and s,
                                  a merged row keeps columns r and s apart) */
#define GSAP_DELTA_EXACT 1     /* the true change of the description length: every entry of
                                  the touched rows and columns with its new degrees, merged
                                  columns folded, self loops moved with the vertex */

/* acceptance of nodal moves */
#define GSAP_ACCEPT_GREEDY 0   /* accept iff delta-S < 0 (kernels.hpp:600-643) */
#define GSAP_ACCEPT_MH 1       /* Metropolis-Hastings: accept with probability
                                  min(1, exp(-beta dS) p(v: s->r) / p(v: r->s)), the proposal
                                  probabilities of the nodal sampler before and after the move */

typedef struct gsap_ctx gsap_ctx;

/* A directed graph in two CSR views (out-adjacency and in-adjacency).  Ids are
 * 0-based int32, weights are positive int32, ptr arrays have N+1 entries (the
 * reference's Csr::adj_ptr has N; the C++ wrapper appends the end).  All
 * pointers are host pointers. */
typedef struct gsap_graph_i32 {
  int64_t N;
  int64_t E;
  const int32_t *out_ptr; /* N+1 */
  const int32_t *out_adj; /* E   */
  const int32_t *out_w;   /* E   */
  const int32_t *in_ptr;  /* N+1 */
  const int32_t *in_adj;  /* E   */
  const int32_t *in_w;    /* E   */
} gsap_graph_i32;

/* Algorithm options.  Defaults (gsap_options_default) are the hyper-parameters
 * hard-coded at GSAP/partition.hpp:855-861. */
typedef struct gsap_options {
  uint64_t seed;              /* Philox key */
  int32_t num_proposals;      /* merge proposals per block (10), 1..32 */
  float reduction_rate;       /* block reduction rate (0.35) */
  int32_t max_nodal_itr;      /* nodal sweeps per level (100) */
  int32_t num_batches;        /* nodal batches per sweep (4) */
  double threshold1;          /* convergence threshold before bracketing (5e-4) */
  double threshold2;          /* convergence threshold after bracketing (1e-4) */
  int32_t moving_avg_window;  /* moving-average window (3) */
  int32_t max_levels;         /* safety cap on outer iterations (0 = none) */
  int32_t precision;          /* GSAP_PRECISION_* */
  uint32_t chain;             /* Philox stream id: distinct chains, distinct streams */
  int32_t trace;              /* 1: record per-step digests (slow, for tests) */
  int32_t dense_max_B;        /* use the dense block matrix when B <= this (0 = auto) */
  int32_t profile;            /* 1: time kernel classes with HIP events (report.kernel_ms) */
  int32_t model;              /* GSAP_MODEL_*: description-length model term */
  int32_t workers;            /* batch / chain calls: concurrent sub-contexts (0 = auto, 8) */
  int32_t profile_classes;    /* with profile = 1: bit k times kernel class k (0 = all) */
  int32_t graph;              /* 1: capture a level's sweep once as a HIP graph and replay it
                                 (same results; ignored with trace or profile); 0: launch
                                 the sweep kernel by kernel */
  int32_t tile_max_B;         /* row-tile rebuild for 128 < B <= this (dense B only;
                                 0 = auto, 2048); above it the segmented rebuild */
  int32_t delta_mode;         /* GSAP_DELTA_* (nodal moves and merges) */
  int32_t accept;             /* GSAP_ACCEPT_* (nodal moves; MH needs precision FP64) */
  double beta;                /* inverse temperature of GSAP_ACCEPT_MH (3.0) */
  /* Work-split choices (results are identical for any valid values; 0 = default).  Nodal
   * moves by a vertex's neighbour count d = max(out, in): d <= nodal_max16 -> 16-lane
   * groups (default 16, at most 16), d <= nodal_max32 -> 32-lane groups (32, at most 32),
   * d <= nodal_maxwave -> one wave (128, at most 128), larger d -> a workgroup; -1 leaves
   * a form unused.  Segmented rebuild: segments of at most seg_reg_max entries are sorted
   * in register groups (64; 16, 32 or 64), longer ones by one wave up to 512 entries; when
   * B <= 16000, segments longer than seg_count_min (512, up to 4096) are counted in LDS
   * instead of sorted. */
  int32_t nodal_max16;
  int32_t nodal_max32;
  int32_t nodal_maxwave;
  int32_t seg_reg_max;
  int32_t seg_count_min;
  /* After a nodal batch at a B without the dense block matrix: 0 (default) rebuilds the
   * block matrix from every edge; 1 merges the batch's moves into the current CSR/CSC as a
   * sorted signed delta (work follows the moved vertices' edges).  Identical results. */
  int32_t sparse_update;
  /* At a B with the dense block matrix: 0 (default) updates only the dense matrix and the
   * block degrees after each nodal batch but a sweep's last, samples the next batch's block
   * neighbours from the dense matrix, and rebuilds the CSR/CSC after the sweep's last batch;
   * 1 rebuilds the CSR/CSC after every batch.  Identical results. */
  int32_t csr_every_batch;
} gsap_options;

/* Per-call report.  Times are seconds of host wall clock around device work. */
typedef struct gsap_report {
  int32_t B;              /* final number of blocks (ids may have gaps) */
  int32_t levels;         /* outer iterations */
  int32_t merge_phases;
  int32_t sweeps;         /* nodal sweeps (4 batches + MDL each) */
  int64_t batches;
  double S;               /* final description length (FP64) */
  double t_total;
  double t_merge;
  double t_sweep;         /* sum over sweeps of sweep wall time */
  double t_rebuild;       /* device time of every block-matrix rebuild (kernel class 1) when
                             options.profile = 1, else 0 */
  int64_t sweep_edges;    /* E * sweeps */
  double max_abs_err_fp32;/* mixed form: max |dS32 - dS64| over proposals evaluated in FP64 */
  int64_t rechecks;       /* mixed form: decisions whose FP32 delta-S lay within its error
                             bound of 0, taken from the FP64 recomputation */
  int64_t n_trace;        /* trace records produced (may exceed capacity) */
  double kernel_ms[8];    /* device time per kernel class when options.profile = 1:     */
  int64_t kernel_calls[8];/* 0 nodal moves, 1 block-matrix rebuild, 2 per-block terms,   */
                          /* 3 merge proposals+delta-S, 4 MDL, 5 accept+batch sum,       */
                          /* 6 nodal proposal pass (unit-weight graphs)                  */
  int64_t fp64_evals;     /* mixed form: nodal proposals evaluated in FP64 (rechecks plus
                             every accepted move, whose exact delta-S enters the batch sum) */
  double t_upload;        /* graph validation, host preparation and upload (resident
                           * graph: the batch lists and workspace only) */
  int64_t nodal_evals;    /* options.profile = 1 on a unit-weight graph: nodal proposals of
                             another block (the moves kernel evaluates these), else -1 */
  int64_t nodal_eval_adj; /* their out+in adjacency entries, summed over batches, else -1 */
} gsap_report;

/* One trace record; identical layout to the oracle's.
 *   kind 1 merge:  B, a = B after merging, b = merge counter, h1 = proposals, h2 = labels
 *   kind 2 batch:  B, a = sweep, b = batch, h1 = proposals, h2 = labels, x = accepted delta-S
 *                  sum, y = accepted moves
 *   kind 3 sweep:  B, a = sweep, h1 = CSR columns, h2 = CSR values, x = S, y = sweep delta-S
 *   kind 4 golden: B, a = next B_to_merge, b = 1 when optimal, h1 = labels, x = S */
typedef struct gsap_trace_rec {
  int32_t kind;
  int32_t B;
  int32_t a, b;
  uint64_t h1, h2;
  double x, y;
} gsap_trace_rec;

void gsap_options_default(gsap_options *o);

int gsap_create(gsap_ctx **ctx, int device);
int gsap_destroy(gsap_ctx *ctx);
const char *gsap_last_error(const gsap_ctx *ctx);

/* Whole partition (GSAP/partition.hpp:852-954).  partitions_out[N] receives
 * 0-based block ids.  report may be NULL. */
int gsap_partition_i32(gsap_ctx *ctx, const gsap_graph_i32 *g, const gsap_options *o,
                       int32_t *partitions_out, gsap_report *report);

/* Device-resident graph.  No counterpart in the reference, whose partition() copies the
 * graph to the device inside the call (partition.hpp:852-870, allocate_device.hpp:4-194);
 * for callers that partition a graph already in device memory.  gsap_upload_graph_i32
 * validates g and copies it into the context (the first step of gsap_partition_i32);
 * gsap_partition_resident_i32 then runs the whole partition on that copy, with no
 * host-to-device transfer of the graph, and gives the result gsap_partition_i32 gives on
 * the same graph and options.  N is the length of partitions_out and must equal the
 * resident graph's vertex count.  Every call on the context that takes a graph replaces
 * the resident copy; without one, gsap_partition_resident_i32 fails with
 * GSAP_ERR_INVALID. */
int gsap_upload_graph_i32(gsap_ctx *ctx, const gsap_graph_i32 *g);
int gsap_partition_resident_i32(gsap_ctx *ctx, const gsap_options *o, int64_t N,
                                int32_t *partitions_out, gsap_report *report);

/* Batched form: n_graphs independent graphs (uneven sizes allowed).  They run on a pool of
 * o->workers sub-contexts, each with its own HIP stream and workspace, so the kernels of
 * different graphs overlap on the device.  outs[k] receives graph k's labels (N_k int32),
 * reports[k] its report (reports may be NULL).  Graph k's result equals
 * gsap_partition_i32 on it alone.  No counterpart in the reference (one graph per run). */
int gsap_partition_batch_i32(gsap_ctx *ctx, const gsap_graph_i32 *graphs, int32_t n_graphs,
                             const gsap_options *o, int32_t *const *outs, gsap_report *reports);

/* Parallel chains: n_chains partitions of one graph with Philox stream ids o->chain + k;
 * out receives the labels of the chain with the lowest S (ties: lowest k), *best_chain its
 * index, S_chain[k] every chain's S (may be NULL).  No counterpart in the reference. */
int gsap_partition_chains_i32(gsap_ctx *ctx, const gsap_graph_i32 *g, const gsap_options *o,
                              int32_t n_chains, int32_t *out, gsap_report *report,
                              double *S_chain, int32_t *best_chain);

/* Autotuner: runs one partition of this graph to take label snapshots at its levels, then
 * times on those labellings, on this device: the row-tile against the segmented rebuild
 * (tile_max_B), dense against CSR block-matrix lookups over a whole batch (dense_max_B),
 * the segment-sort classes (seg_reg_max, seg_count_min) and the nodal neighbour-count
 * classes (nodal_max16, nodal_max32, nodal_maxwave), each weighted by the batches the
 * partition ran at that level, and writes the fastest choices into *o (the other fields
 * are kept).  These fields only choose code paths: results are identical for any values.
 * No counterpart in the reference (its bins are fixed, bb_segsort/bb_bin.h:23). */
int gsap_autotune_i32(gsap_ctx *ctx, const gsap_graph_i32 *g, gsap_options *o);

/* Trace records of the last gsap_partition_i32 call with o->trace = 1. */
int gsap_get_trace(const gsap_ctx *ctx, gsap_trace_rec *out, int64_t cap, int64_t *n);

/* Block matrix of a labelling (GSAP/partition.hpp:707-849).  Caller-allocated
 * outputs: rowptr[B+1], col[E], val[E], colptr[B+1], row[E], cval[E], dout[B],
 * din[B].  *nnz receives the number of nonzero blocks pairs. */
int gsap_blockmodel_i32(gsap_ctx *ctx, const gsap_graph_i32 *g, const int32_t *labels, int32_t B,
                        int64_t *rowptr, int32_t *col, int64_t *val, int64_t *colptr, int32_t *row,
                        int64_t *cval, int64_t *dout, int64_t *din, int64_t *nnz,
                        const gsap_options *o);

/* One block-merge phase at merge counter m (GSAP/partition.hpp:324-454, 26-78).
 * labels[N] is updated in place; prop[np*B], dS[np*B], best[B], dS_best[B] may
 * be NULL.  *B_new receives the block count after merging. */
int gsap_merge_phase_i32(gsap_ctx *ctx, const gsap_graph_i32 *g, const gsap_options *o,
                         int32_t *labels, int32_t B, int32_t B_to_merge, uint32_t m, int32_t *prop,
                         double *dS, int32_t *best, double *dS_best, int32_t *B_new);

/* One nodal batch b at batch counter q (GSAP/partition.hpp:456-662).  labels[N]
 * updated in place; prop[BS], dS[BS] may be NULL; *dE receives the sum of the
 * accepted delta-S. */
int gsap_nodal_batch_i32(gsap_ctx *ctx, const gsap_graph_i32 *g, const gsap_options *o,
                         int32_t *labels, int32_t B, int32_t b, uint32_t q, int32_t *prop,
                         double *dS, double *dE);

/* One nodal batch (as gsap_nodal_batch_i32) followed by the block-matrix update that the
 * partition loop runs after every batch (calculate_new_csr, GSAP/partition.hpp:914-915):
 * the dense-matrix update, the sparse merge of the batch's moves or a full rebuild, as the
 * options and B select.  keep = 0: starts from labels[N] (block matrix rebuilt); keep = 1:
 * continues from the labels and block matrix the previous call on this context left (same
 * graph, B and num_batches).  labels[N] receives the labels after the batch; the block
 * matrix afterwards is written like gsap_blockmodel_i32. */
int gsap_nodal_batch_bm_i32(gsap_ctx *ctx, const gsap_graph_i32 *g, const gsap_options *o,
                            int32_t *labels, int32_t B, int32_t b, uint32_t q, int32_t keep,
                            int64_t *rowptr, int32_t *col, int64_t *val, int64_t *colptr,
                            int32_t *row, int64_t *cval, int64_t *dout, int64_t *din,
                            int64_t *nnz);

/* Description length of a labelling (GSAP/partition.hpp:665-704), FP64. */
int gsap_mdl_i32(gsap_ctx *ctx, const gsap_graph_i32 *g, const int32_t *labels, int32_t B,
                 double *S, const gsap_options *o);

/* Description length of a labelling without allocating device memory.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Every input and the result are the
 * caller's DEVICE buffers: d_out_ptr[N+1], d_out_adj[E], d_out_w[E] (the out-CSR view),
 * d_labels[N]; d_S receives the FP64 description length, equal bit for bit to what
 * gsap_mdl_i32 returns for the same graph and labels (NaN when an id, weight or row pointer
 * is out of range).  Limit: 1 <= B <= 128 (GSAP_ERR_INVALID otherwise).  Work is queued on
 * `stream` (a hipStream_t; NULL = the context's stream) and not waited for.  Intermediates
 * live in one fixed device array of about 4.5 MiB per GPU, held until the process ends and
 * shared by every call on that GPU: calls that overlap on one GPU race. */
int gsap_mdl_noalloc_i32(gsap_ctx *ctx, const int32_t *d_out_ptr, const int32_t *d_out_adj,
                         const int32_t *d_out_w, const int32_t *d_labels, int64_t N, int64_t E,
                         int32_t B, int32_t model, double *d_S, void *stream);

/* ---- host-side helpers (no device needed) ---- */

/* Golden-section search over the block count (host only).  A slot is the reference's
 * Partition (GSAP/graph.hpp:48-54) with S in FP64 and the id (0..2) of the label snapshot
 * it owns; the three slots are OldData's large, med, small (graph.hpp:57-70). */
typedef struct gsap_golden_slot {
  int32_t B;
  int32_t B_to_merge;
  double S;
  int32_t snap;
  int32_t reserved;
} gsap_golden_slot;

/* slots[k] = {B 0, S inf, snapshot k}: OldData() of GSAP/graph.hpp:62-69. */
void gsap_golden_init(gsap_golden_slot slots[3]);

/* Convergence test after sweep itr (GSAP/partition.hpp:918-934): returns 1 when the level
 * stops.  dE[0..itr] are the sweeps' accepted delta-S sums, W the moving-average window,
 * bracketed = 1 when all three slots hold a finite S (threshold2 then, else threshold1). */
int32_t gsap_converged(const double *dE, int32_t itr, int32_t W, double S, int32_t bracketed,
                       double threshold1, double threshold2);

/* Bracket update after a level (prepare_for_partition_next, GSAP/partition.hpp:80-322).
 * slots = {large, med, small}, updated in place; P = the level's result (B, B_to_merge, S).
 * *store receives the snapshot that must take the level's labels, *restore the snapshot the
 * next level starts from, *next its B, the next B_to_merge and its S, *optimal 1 when the
 * search ends.  One addition to the reference: a step of B_to_merge <= 0 in the bracketed
 * phase also ends the search (the reference would repeat the level without merging). */
void gsap_golden_step(gsap_golden_slot slots[3], const gsap_golden_slot *P, float rate,
                      int32_t *store, int32_t *restore, gsap_golden_slot *next,
                      int32_t *optimal);

/* Seeded degree-corrected SBM.  Two-phase: call with out arrays NULL to learn
 * E (= llround(N*mean_deg)), then again with arrays of the right sizes.
 * alpha <= 0 gives near-equal block sizes, else Dirichlet(alpha) sizes. */
typedef struct gsap_dcsbm_params {
  int64_t N;
  int32_t K;
  double alpha;
  double rho;       /* intra / inter edge ratio */
  double mean_deg;
  int32_t deg_min;  /* 5 */
  int32_t deg_max;  /* 100 */
  double exponent;  /* 2.5 */
  uint64_t seed;
} gsap_dcsbm_params;

int gsap_generate_dcsbm_i32(const gsap_dcsbm_params *p, int64_t *E, int32_t *out_ptr,
                            int32_t *out_adj, int32_t *out_w, int32_t *in_ptr, int32_t *in_adj,
                            int32_t *in_w, int32_t *truth);

/* GraphChallenge TSV (1-based "src dst w" lines).  Two-phase like the
 * generator: first call with arrays NULL to get N and E. */
int gsap_read_tsv_i32(const char *path, int64_t *N, int64_t *E, int32_t *out_ptr, int32_t *out_adj,
                      int32_t *out_w, int32_t *in_ptr, int32_t *in_adj, int32_t *in_w);

/* Binary CSR cache: the graph's two CSR views (and optionally a truth labelling) with a
 * checksum, so large graphs reload without parsing or regeneration.  Load is two-phase:
 * first with out_ptr NULL (N, E and has_truth only), then with arrays of the right sizes
 * (truth[N] needed when has_truth).  GSAP_ERR_IO for a missing, truncated, foreign or
 * corrupted file.  No counterpart in the reference. */
int gsap_save_csr_i32(const char *path, const gsap_graph_i32 *g, const int32_t *truth);
int gsap_load_csr_i32(const char *path, int64_t *N, int64_t *E, int32_t *has_truth,
                      int32_t *out_ptr, int32_t *out_adj, int32_t *out_w, int32_t *in_ptr,
                      int32_t *in_adj, int32_t *in_w, int32_t *truth);

/* Pairwise precision / recall against a truth labelling (int64 counts). */
int gsap_pairwise_i32(int64_t N, const int32_t *truth, const int32_t *given, double *precision,
                      double *recall);

/* Full quality report against a truth labelling (the metrics of bf::evaluate,
 * GSAP/evaluate/evaluate.hpp:19-265, with int64 pair counts and FP64 sums; vertices
 * whose truth is negative are left out).  accuracy is the best one-to-one matching of
 * truth and given blocks (Kuhn-Munkres) over n; computed (accuracy_computed = 1) when the
 * contingency table has at most 64M cells.  nmi = 2 |MI| / (|H_truth| + |H_given|). */
typedef struct gsap_quality {
  int64_t n;                 /* vertices with a truth label */
  int32_t B_truth;           /* nonempty truth blocks */
  int32_t B_given;           /* max given label + 1 */
  double accuracy;
  double precision;          /* pairwise */
  double recall;             /* pairwise */
  double rand_index;
  double adjusted_rand_index;
  double nmi;
  int32_t accuracy_computed;
  int32_t reserved;
} gsap_quality;

int gsap_quality_i32(int64_t N, const int32_t *truth, const int32_t *given, gsap_quality *q);

/* Canonical log2 used by every FP64 delta-S / MDL term (host evaluation). */
double gsap_log2(double x);
/* Its constants: inv_c[128], log2_c[128], poly[9]. */
void gsap_log2_tables(double *inv_c, double *log2_c, double *poly);
/* FP32 log2 of the mixed form (host evaluation of the device code), y[k] = log2(x[k]) for
 * x[k] >= 1. */
void gsap_log2f_n(const float *x, float *y, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* GSAP_GSAP_H */
