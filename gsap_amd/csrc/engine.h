// Device state of one partition run and the kernel entry points the host
// orchestrator (engine.hip) drives.
#pragma once
#include <string>
#include <utility>

#include "common.h"
#include "prims.h"

namespace gsap {

// Graph in HBM: out-CSR, in-CSR (N+1 pointers), weighted degrees and the COO source
// of every out-edge (so edge-parallel kernels read it coalesced).
struct GraphDev {
  int64_t N = 0, E = 0;
  int32_t *optr = nullptr, *oadj = nullptr, *ow = nullptr;
  int32_t *iptr = nullptr, *iadj = nullptr, *iw = nullptr;
  int32_t *kout = nullptr, *kin = nullptr, *esrc = nullptr;
  int32_t *idst = nullptr;  // owner (target vertex) of every in-CSR entry
  int32_t *sloop = nullptr; // per vertex: summed weight of its self loops
  int32_t wmax = 1;          // largest edge weight
  // nodal hash tables for vertices whose adjacency does not fit the LDS table
  int64_t *hoff = nullptr;       // per vertex: offset into hpool (words), -1 when light
  uint8_t *hbits_o = nullptr;    // per vertex: log2 capacity of out table (0 = light)
  uint8_t *hbits_i = nullptr;
  int32_t *hpool = nullptr;      // 3 words per slot, out table then in table
};

// Block matrix of the current labelling.  rowptr[B] holds nnz on the device.
struct BMBuf {
  int32_t B = 0;
  int32_t *rowptr = nullptr, *colptr = nullptr;   // Bmax+1
  int32_t *col = nullptr, *val = nullptr, *erow = nullptr;  // E
  int32_t *row = nullptr, *cval = nullptr;        // E
  int32_t *dout = nullptr, *din = nullptr;        // Bmax
  double *ldout = nullptr, *ldin = nullptr;       // Bmax
  int32_t *dense = nullptr;                       // dense_cap words
  int64_t dense_cap = 0;
  int32_t *gcol = nullptr, *grow = nullptr;       // E/16 + 1: coarse index of col / row
  bool index_valid = false;
  bool dense_valid = false;

  // the transpose of the dense matrix in the rebuild scratch's partial-plane buffer is
  // current for B = dt_B (0: not current); the incremental update keeps it so
  int32_t dt_B = 0;
  // the dense matrix's column chunk sums (rebuild scratch colsum) are current for B = cs_B
  int32_t cs_B = 0;
  // B of the entries the dense buffer holds (the current CSR's); 0 = all zero
  int32_t dense_B = 0;
  // the dense matrix and the block degrees are current but the CSR/CSC (and nnz in
  // rowptr[B]) and the log2 degrees still hold an earlier labelling's matrix: set by a
  // deferred update between the batches of a sweep, cleared by every full CSR build
  bool csr_stale = false;
  // The alternate CSR/CSC buffers (same sizes): the sparse update after a nodal batch merges
  // the current matrix and the batch's delta into them, then the two sets change roles;
  // flipped records that the alternate set holds the context's primary buffers.
  int32_t *a_rowptr = nullptr, *a_colptr = nullptr, *a_col = nullptr, *a_val = nullptr,
          *a_erow = nullptr, *a_row = nullptr, *a_cval = nullptr, *a_gcol = nullptr,
          *a_grow = nullptr;
  bool flipped = false;
  void swap_sets() {
    std::swap(rowptr, a_rowptr); std::swap(colptr, a_colptr); std::swap(col, a_col);
    std::swap(val, a_val); std::swap(erow, a_erow); std::swap(row, a_row);
    std::swap(cval, a_cval); std::swap(gcol, a_gcol); std::swap(grow, a_grow);
    flipped = !flipped;
  }
  BMView view() const {
    BMView v;
    v.B = B; v.rowptr = rowptr; v.col = col; v.val = val;
    v.colptr = colptr; v.row = row; v.cval = cval;
    v.dout = dout; v.din = din; v.ldout = ldout; v.ldin = ldin;
    v.dense = dense_valid ? dense : nullptr;
    v.gcol = index_valid ? gcol : nullptr;
    v.grow = index_valid ? grow : nullptr;
    return v;
  }
};

constexpr int SEG_NCL = 10;  // segment classes of the segmented rebuild

struct RebuildScratch {
  void *k0 = nullptr, *k1 = nullptr;  // 2E words each: segment keys and weights
  uint32_t *tmp = nullptr;            // scan temporary
  int64_t tmp_words = 0;
  int32_t *partials = nullptr;        // partial planes of the small-B histogram
  int64_t partials_cap = 0;           // words
  // above 8192 blocks (dense): colsum[x * ceil(B / 64) + c] = sum of column x of the dense
  // matrix over rows 64 c .. 64 c + 63 (the column walks' index; no transpose is kept there)
  int32_t *colsum = nullptr;
  int64_t colsum_cap = 0;             // words
  uint32_t *soff = nullptr;           // 2 Bmax + 1: segment sizes, then offsets
  uint32_t *su = nullptr;             // 2 Bmax: runs per segment
  uint32_t *vrel = nullptr;           // 2N: each vertex's slice start inside its row / column segment
  uint32_t *chunk_seg = nullptr;      // 2E / 64 + 1: segment holding each 64-entry chunk's start
  uint16_t *lab16 = nullptr;          // N: 16-bit labels when B <= 65535
  uint32_t *suoff = nullptr;          // 2 Bmax + 1: offsets of the runs
  uint32_t *lists = nullptr;          // SEG_NCL lists of lcap segment ids
  int64_t lcap = 0;
  uint32_t *ncl = nullptr;            // SEG_NCL list lengths
  uint32_t *cntc = nullptr;           // runs per (large segment, column range)
  // segment classes: at most seg_reg_max entries sorted in register groups (16, 32 or 64);
  // with B in one column range, segments longer than seg_count_min (512..4096) counted in LDS
  int32_t seg_reg_max = 64, seg_count_min = 512;
  // sparse update after a nodal batch: k0 / k1 hold kcap words each; rk / rw (rcap words
  // each) receive the runs of segments longer than 4096 entries when B > 16000
  int64_t kcap = 0, rcap = 0;
  uint32_t *rk = nullptr, *rw = nullptr;
  bool sparse_on = true;
  // its work area: the moved list (mwork_list words), then 8 counter words
  uint32_t *mwork = nullptr;
  int64_t mwork_list = 0;
};
// words of cntc for a graph of N vertices and E edges
int64_t seg_cntc_words(int64_t N, int64_t E);

// ---- bm.hip
void rebuild_blockmodel(const GraphDev &g, const int32_t *lab, int32_t B, BMBuf &bm,
                        RebuildScratch &rs, int32_t dense_max_B, int32_t tile_max_B,
                        hipStream_t st);

// After a nodal batch over [lo, lo + BS) (olab: the batch's labels before it), the block
// matrix by moving the moved vertices' edges between cells of the dense matrix, when that
// holds the matrix of the labels before the batch (dense B, previous rebuild at this B).
// Returns false (nothing launched) otherwise (dense_update_ok).  defer_csr (the batch's
// apply has moved the block degrees already): only the dense matrix (its transpose too above
// 512 blocks, kept in the partial-plane buffer with bm.dt_B = B) follows the batch; the CSR,
// CSC and log2 degrees are left for a later rebuild (bm.csr_stale).
bool dense_update_ok(const BMBuf &bm, const RebuildScratch &rs, int32_t B);
bool rebuild_after_moves(const GraphDev &g, const int32_t *lab32, const int32_t *olab, int64_t lo,
                         int64_t BS, int32_t B, BMBuf &bm, RebuildScratch &rs, hipStream_t st,
                         bool defer_csr = false);

// After a nodal batch over [lo, lo + BS) at a B without the dense copy: the batch's edge
// moves as a signed delta matrix, sorted per row and column by the segmented rebuild's
// sorts and merged with the current CSR and CSC into the alternate buffers (then swapped
// in); bound = an upper bound of the delta entries (4 x the batch's out- plus in-adjacency).
// Returns false (nothing launched) when the current matrix is not the batch-start one at
// this B or the delta does not fit the scratch.
bool rebuild_after_moves_sparse(const GraphDev &g, const int32_t *lab32, const int32_t *olab,
                                int64_t lo, int64_t BS, int32_t B, int64_t bound, BMBuf &bm,
                                RebuildScratch &rs, hipStream_t st);

// D = sum of C partial planes of B*B counters, then the block matrix of D (B <= 128) as CSR,
// CSC, degrees and log2 degrees (the small rebuild path's reduction and compaction)
void launch_small_from_planes(const int32_t *partials, int32_t C, int32_t B, int32_t *D,
                              int32_t *rowptr, int32_t *col, int32_t *val, int32_t *erow,
                              int32_t *colptr, int32_t *row, int32_t *cval, int32_t *dout,
                              int32_t *din, double *ldout, double *ldin, hipStream_t st);

// ---- noalloc.hip
// Resolves this device's address of the no-allocation entry point's scratch array (called
// once per device with the device current, by gsap_create); returns false on failure.
bool noalloc_init_device(int device);
// Description length of a labelling (B <= 128) on device buffers with every intermediate in
// the scratch array; *S (device) receives it.  Returns a GSAP_* status, msg on error.
int noalloc_mdl(int device, const int32_t *optr, const int32_t *oadj, const int32_t *ow,
                const int32_t *lab, int64_t N, int64_t E, int32_t B, int32_t model_kind,
                double *S, hipStream_t st, std::string &msg);

// ---- sweep.hip
struct RngKey {
  uint32_t k0, k1, chain;
};
// Philox phase counter of a nodal batch: *dev + add (dev may be null).  Kept in device
// memory so that a captured sweep replays with the current batch counter.
struct PhaseRef {
  const uint32_t *dev;
  uint32_t add;
};
// zero: NCLS words cleared by the launch (the nodal proposal pass's counters), or null
void launch_block_terms(const BMView &m, double *OUT, int32_t *sof, uint8_t *coin, RngKey key,
                        PhaseRef phase, bool sample, int32_t avg_row, hipStream_t st,
                        int32_t *zero = nullptr);
// The per-block samples and coins from the dense matrix D (B x B, row-major) and, when not
// null, its transpose DT, for batches whose CSR/CSC lags the labels; also writes the log2
// block degrees of dout / din.
// (without DT, CS: the column chunk sums of D, 64 rows per chunk, for the column walks)
void launch_block_terms_dense(const int32_t *D, const int32_t *DT, const int32_t *CS, int32_t B,
                              const int32_t *dout, const int32_t *din, double *ldout,
                              double *ldin, int32_t *sof, uint8_t *coin, RngKey key,
                              PhaseRef phase, hipStream_t st, int32_t *zero);
// Nodal kernels by the larger of a vertex's out- and in-neighbour counts d: d <= 16 and 32
// use register deduplication in groups of 16 and 32 lanes; d <= 128 a wave with 256-slot
// LDS hash tables; larger d a 256-thread workgroup with tables sized for the vertex (LDS
// up to big_lds bytes, else its slice of the global pool).
constexpr int NCLS = 4;
// Largest neighbour count of each form (at most 16, 32 and 128; -1: the form unused)
struct NodalBounds {
  int32_t max16 = 16, max32 = 32, maxwave = 128;
  bool operator==(const NodalBounds &o) const {
    return max16 == o.max16 && max32 == o.max32 && maxwave == o.maxwave;
  }
};
inline int nodal_class(int64_t d, const NodalBounds &nb = NodalBounds()) {
  return d <= nb.max16 ? 0 : d <= nb.max32 ? 1 : d <= nb.maxwave ? 2 : 3;
}
// Space (int32 words) a class-3 vertex needs: three int32 arrays per hash table of 2^bits
// slots (bits = smallest with 2^bits >= 2 n) and one double per row and column slot
inline int64_t big_table_bits(int64_t n) {
  int64_t b = 1;
  while (((int64_t)1 << b) < 2 * (n > 1 ? n : 1)) ++b;
  return b;
}
inline int64_t big_words(int64_t no, int64_t ni) {
  return 3 * (((int64_t)1 << big_table_bits(no)) + ((int64_t)1 << big_table_bits(ni))) +
         2 * (no + 2 + ni);
}
constexpr int64_t BIG_LDS_CAP = 40 * 1024;
// Unit-weight graphs: the proposal pass lists the vertices that propose another block by
// class (dlist: NCLS x BS records) and counts them in dcnt (NCLS words, cleared by the batch's
// block terms before it).  dlist = null: no proposal pass.
// evstat, or null: running totals (vertices evaluated, their adjacency entries) in EV_SLOTS
// pairs, which the host sums.
constexpr int EV_SLOTS = 32;
struct NodalPre {
  int4 *dlist = nullptr;
  int32_t *dcnt = nullptr;
  unsigned long long *evstat = nullptr;
  NodalBounds bounds;
};
// stages of launch_nodal: the proposal pass (unit-weight graphs) and the moves kernel
constexpr int NODAL_PROPOSE = 1, NODAL_MOVES = 2;
void launch_nodal(const GraphDev &g, LabView lab, const BMView &m, const int32_t *sof,
                  const uint8_t *coin, RngKey key, PhaseRef phase, int64_t lo, int64_t BS,
                  const int32_t *const lists[NCLS], const int32_t counts[NCLS], uint32_t big_lds,
                  int32_t *prop, double *dS, float *err, uint8_t *near, bool exact, bool mh,
                  double beta, const NodalPre &pre, int stages, hipStream_t st);
void launch_mix_stats(const float *err, const uint8_t *near, int64_t n, uint32_t *max_bits,
                      unsigned long long *count, hipStream_t st);
// Labels of a nodal batch (greedy, all = false: where dS < 0; Metropolis-Hastings, all = true:
// prop holds every vertex's new label; olab keeps the old ones, lab16 when given follows)
// and the canonical sum of the batch's dS into *out (out = nullptr: tile sums only, the
// total left to launch_sweep_sums).
// With dm.dout set (a batch whose CSR rebuild is deferred), a moved vertex's weighted out-
// and in-degree also leave its old block and enter its new one in dm.dout / dm.din.
struct DegMoves {
  const int32_t *optr = nullptr, *iptr = nullptr;  // adjacency counts (unit weights), or
  const int32_t *kout = nullptr, *kin = nullptr;   // weighted degrees
  int32_t B = 0;
  int32_t *dout = nullptr, *din = nullptr;
};
void launch_apply_sum(int32_t *lab, const int32_t *prop, const double *dS, int64_t lo, int64_t BS,
                      bool all, int32_t *olab, uint16_t *lab16, double *tiles, double *out,
                      const DegMoves &dm, hipStream_t st);
// The end of a sweep: the description-length data term (tiles at tiles[0..)) into out[0] and
// the deferred delta-S sums of the nb batches (tiles at tiles + boff + b * bstride) into
// out[1 + b], one launch; ctr advanced by add afterwards.
void launch_sweep_sums(const BMView &m, const int32_t *erow, int64_t n_max, double *tiles,
                       int64_t boff, int64_t bstride, int32_t nb, int64_t BS, double *out,
                       uint32_t *ctr, uint32_t add, hipStream_t st);
// res[0..n) and *nnz copied into host memory mapped for the device (hv, hnnz), then seq
// stored at *hseq with system scope once they are visible to the host
void launch_publish(const double *res, int32_t n, const int32_t *nnz, double *hv, int32_t *hnnz,
                    uint32_t *hseq, uint32_t seq, hipStream_t st);
// description-length data term summed in canonical order into *out; ctr (when given)
// advanced by add afterwards
void launch_mdl_sum(const BMView &m, const int32_t *erow, int64_t n_max, double *tiles,
                    double *out, uint32_t *ctr, uint32_t add, hipStream_t st);
// Merge proposals and delta-S of every block.  big_list (B words) and big_count (one word)
// serve the workgroup-per-block form at large B (it synchronizes the stream once); without
// them every proposal gets a group of lanes.
void launch_merge(const BMView &m, const double *OUT, RngKey key, uint32_t phase, int32_t np,
                  int32_t *prop, double *dS, int64_t nnz, bool exact, int32_t *big_list,
                  uint32_t *big_count, hipStream_t st);
void launch_argmin(const double *dS, const int32_t *prop, int32_t B, int32_t np, int32_t *best,
                   double *dS_best, hipStream_t st);
// After a merge phase: par[B] is the merge forest (par[x] == x for the surviving blocks).
// Every block's pointer is replaced by its root, the roots are numbered in increasing id
// order (rank, B words; tmp: scan_temp_words(B)), and lab[v] becomes its block's root rank.
void launch_merge_relabel(int32_t *par, int32_t B, uint32_t *rank, uint32_t *tmp, int32_t *lab,
                          int64_t N, hipStream_t st);
void launch_iota(int32_t *x, int64_t n, hipStream_t st);
// owner[j] = v for every entry j of vertex v's row of the CSR ptr[N + 1]
void launch_owners(const int32_t *ptr, int64_t N, int32_t *owner, hipStream_t st);
// keys[r] = order key of the double x[r], vals[r] = r
void launch_order_keys(const double *x, int32_t n, uint64_t *keys, uint32_t *vals,
                       hipStream_t st);

}  // namespace gsap
