// Block-matrix rebuild: labels -> M as CSR (columns ascending) + CSC (rows ascending),
// degrees and cached log2 of the degrees.  Replaces calculate_new_csr,
// GSAP/partition.hpp:707-849 (sort_by_key of vertices, node-level gathers, two bb_segsort
// calls, run flags, remove_if, scans and mgpu::segreduce).
//
// Two paths, chosen by B:
//   small    (B <= 128 with the dense matrix): the whole B x B histogram in LDS per
//            workgroup over a chunk of the edges, a reduction of the partial planes, then
//            CSR + CSC + degrees from D in one workgroup.
//   segmented (every other B): segment x < B holds the out-edges of the vertices of block
//            x keyed by the target's block (row x of M), segment B + x the in-edges of
//            block x keyed by the source's block (column x).  Vertices reserve their slice
//            of a segment with one atomic; each segment is then sorted and run-length
//            reduced by a team sized to it (registers for <= 64 entries, one wave's LDS for
//            <= 512, one workgroup's LDS for <= 4096, LDS counters over column ranges
//            above that).  The integer results do not depend on the order in which the
//            vertices reserved their slices.
// The dense B x B copy of M (kept when B <= dense_max_B, for O(1) lookups by the sweep
// kernels) is maintained in O(nnz): every rebuild first clears the previous nonzeros, so
// the buffer holds zeros everywhere else.
#include <algorithm>

#include "engine.h"

namespace gsap {

// dense[r * B + c] = 0 for every entry (r, c) of the previous CSR (nnz on the device)
__global__ void k_dense_clear(const int32_t *erow, const int32_t *col, const int32_t *nnz_dev,
                              int32_t B, int32_t *dense) {
  const int32_t nnz = *nnz_dev;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x)
    dense[(int64_t)erow[k] * B + col[k]] = 0;
}

__global__ void k_lab16(const int32_t *lab, int64_t N, uint16_t *lab16) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x)
    lab16[v] = (uint16_t)lab[v];
}

// ------------------------------------------------------------------ small path
// B <= 128: the whole B x B histogram in LDS; workgroup c takes a contiguous chunk of the
// edges (id order, one edge per thread per step, so that many label gathers are in flight)
// and writes its plane of partial counts (or D itself when C == 1).
__global__ __launch_bounds__(512) void k_edge_hist(const int32_t *esrc, const int32_t *oadj,
                                                   const int32_t *ow, LabView lab,
                                                   int64_t E, int32_t B, int32_t C,
                                                   int32_t *partial) {
  extern __shared__ int32_t T[];
  const int32_t nt = B * B;
  for (int32_t i = threadIdx.x; i < nt; i += blockDim.x) T[i] = 0;
  __syncthreads();
  const int64_t c = blockIdx.x;
  const int64_t a0 = E * c / C, a1 = E * (c + 1) / C;
  for (int64_t e = a0 + threadIdx.x; e < a1; e += blockDim.x)
    atomicAdd(&T[lab[esrc[e]] * B + lab[oadj[e]]], ow[e]);
  __syncthreads();
  int32_t *out = partial + c * (int64_t)nt;
  for (int32_t i = threadIdx.x; i < nt; i += blockDim.x) out[i] = T[i];
}

// D[i] = sum over the C partial planes; a group of L lanes per counter (L = 1 when C is
// small, up to 64 when a few counters have many planes)
template <int L>
__global__ __launch_bounds__(256) void k_hist_reduce(const int32_t *partials, int32_t nparts,
                                                     int64_t BB, int32_t *D) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / L;
  const int q = threadIdx.x & (L - 1);
  if (i >= BB) return;
  int32_t s = 0;
  for (int32_t p = q; p < nparts; p += L) s += partials[(int64_t)p * BB + i];
#pragma unroll
  for (int off = L / 2; off >= 1; off >>= 1) s += __shfl_down(s, off, L);
  if (q == 0) D[i] = s;
}

__global__ void k_logs(const int32_t *dout, const int32_t *din, int32_t B, double *ldout,
                       double *ldin) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < B;
       x += (int64_t)gridDim.x * blockDim.x) {
    ldout[x] = clog2((double)dout[x]);
    ldin[x] = clog2((double)din[x]);
  }
}

// Unit weights: block degrees are the row and column segment sizes (off: the scanned
// segment offsets, 2B + 1), then their logs.
__global__ void k_logs_seg(const uint32_t *off, int32_t B, int32_t *dout, int32_t *din,
                           double *ldout, double *ldin) {
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < B;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int32_t o = (int32_t)(off[x + 1] - off[x]), i = (int32_t)(off[B + x + 1] - off[B + x]);
    dout[x] = o;
    din[x] = i;
    ldout[x] = clog2((double)o);
    ldin[x] = clog2((double)i);
  }
}

// Small block matrices (B <= 128): CSR, CSC, degrees and their logs from D in one
// workgroup of 1024 threads (16 waves).  Row r (column c) is one wave's ballot compaction
// of D's row (column) in ascending order.
constexpr int SMALL_B = 128;
__global__ __launch_bounds__(1024) void k_small_bm(const int32_t *D, int32_t B, int32_t *rowptr,
                                                   int32_t *col, int32_t *val, int32_t *erow,
                                                   int32_t *colptr, int32_t *row, int32_t *cval,
                                                   int32_t *dout, int32_t *din, double *ldout,
                                                   double *ldin) {
  __shared__ int32_t rcnt[SMALL_B], ccnt[SMALL_B], roff[SMALL_B + 1], coff[SMALL_B + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // counts and sums: waves 0..15 take rows, then columns
  for (int32_t x = w; x < 2 * B; x += 16) {
    const bool is_row = x < B;
    const int32_t y = is_row ? x : x - B;
    uint32_t n = 0;
    int32_t sum = 0;
    for (int32_t t = lane; t < B; t += 64) {
      const int32_t d = is_row ? D[y * B + t] : D[t * B + y];
      n += d != 0;
      sum += d;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      n += __shfl_down(n, off, 64);
      sum += __shfl_down(sum, off, 64);
    }
    if (lane == 0) {
      if (is_row) { rcnt[y] = (int32_t)n; dout[y] = sum; ldout[y] = clog2((double)sum); }
      else { ccnt[y] = (int32_t)n; din[y] = sum; ldin[y] = clog2((double)sum); }
    }
  }
  __syncthreads();
  // exclusive scans of the counts (B <= 128: two entries per lane)
  if (w < 2) {
    const int32_t *cnt = w == 0 ? rcnt : ccnt;
    int32_t *off = w == 0 ? roff : coff;
    const int32_t a0 = 2 * lane < B ? cnt[2 * lane] : 0;
    const int32_t a1 = 2 * lane + 1 < B ? cnt[2 * lane + 1] : 0;
    int32_t incl = a0 + a1;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int32_t yv = __shfl_up(incl, d, 64);
      if (lane >= d) incl += yv;
    }
    const int32_t ex = incl - (a0 + a1);
    if (2 * lane < B) off[2 * lane] = ex;
    if (2 * lane + 1 < B) off[2 * lane + 1] = ex + a0;
    if (lane == 63) off[B] = incl;
  }
  __syncthreads();
  for (int32_t x = threadIdx.x; x <= B; x += 1024) {
    rowptr[x] = roff[x];
    colptr[x] = coff[x];
  }
  // entries
  for (int32_t x = w; x < 2 * B; x += 16) {
    const bool is_row = x < B;
    const int32_t y = is_row ? x : x - B;
    int32_t run = is_row ? roff[y] : coff[y];
    for (int32_t base = 0; base < B; base += 64) {
      const int32_t t = base + lane;
      const int32_t d = t < B ? (is_row ? D[y * B + t] : D[t * B + y]) : 0;
      const uint64_t m = __ballot(d != 0);
      if (d != 0) {
        const int32_t pos = run + (int32_t)__popcll(m & lt);
        if (is_row) { col[pos] = t; val[pos] = d; if (erow) erow[pos] = y; }
        else { row[pos] = t; cval[pos] = d; }
      }
      run += (int32_t)__popcll(m);
    }
  }
}

// ------------------------------------------------------------------- tile path
// ---- vertices grouped by block (counting sort; order inside a block is irrelevant)
constexpr int VT = 2048;  // vertices per workgroup
constexpr int TILE_MAX_B = 8192;  // largest B of the row-tile path
constexpr int INC_MAX_B = 24576;  // largest B of the incremental dense update
constexpr int TAIL_MAX_B = 8192;  // above it the CSR after an update comes from the edges
constexpr int INC_TILE_B = 512;   // its LDS row-tile form runs up to this B, global atomics above

__global__ __launch_bounds__(256) void k_label_hist(LabView lab, int64_t N, int32_t B,
                                                    uint32_t *H, int32_t W) {
  extern __shared__ uint32_t cnt[];
  for (int32_t b = threadIdx.x; b < B; b += blockDim.x) cnt[b] = 0;
  __syncthreads();
  const int64_t v0 = (int64_t)blockIdx.x * VT;
  for (int64_t v = v0 + threadIdx.x; v < v0 + VT && v < N; v += blockDim.x)
    atomicAdd(&cnt[lab[v]], 1u);
  __syncthreads();
  for (int32_t b = threadIdx.x; b < B; b += blockDim.x) H[(int64_t)b * W + blockIdx.x] = cnt[b];
}

__global__ __launch_bounds__(256) void k_label_scatter(LabView lab, int64_t N, int32_t B,
                                                       const uint32_t *H, int32_t W,
                                                       int32_t *vperm) {
  extern __shared__ uint32_t cur[];
  for (int32_t b = threadIdx.x; b < B; b += blockDim.x) cur[b] = H[(int64_t)b * W + blockIdx.x];
  __syncthreads();
  const int64_t v0 = (int64_t)blockIdx.x * VT;
  for (int64_t v = v0 + threadIdx.x; v < v0 + VT && v < N; v += blockDim.x)
    vperm[atomicAdd(&cur[lab[v]], 1u)] = (int32_t)v;
}

// ---- M by row tiles in LDS: workgroup (tile, chunk) owns rows [r0, r0+R) and one chunk of
// those rows' vertices; each wave takes a vertex and its lanes add its out-edges into the
// LDS tile; the tile goes out as a plane of partial counts (no global atomics).
__global__ __launch_bounds__(512) void k_tile_hist(const int32_t *optr, const int32_t *oadj,
                                                   const int32_t *ow, LabView lab,
                                                   const int32_t *vperm, const uint32_t *H,
                                                   int32_t W, int64_t N, int32_t B, int32_t R,
                                                   int32_t C, int32_t *partial) {
  extern __shared__ int32_t T[];
  const int32_t r0 = blockIdx.x * R;
  const int32_t r1 = r0 + R < B ? r0 + R : B;
  const int32_t nt = (r1 - r0) * B;
  for (int32_t i = threadIdx.x; i < nt; i += blockDim.x) T[i] = 0;
  __syncthreads();
  const int64_t vs = H[(int64_t)r0 * W];
  const int64_t ve = r1 < B ? (int64_t)H[(int64_t)r1 * W] : N;
  const int64_t c = blockIdx.y;
  const int64_t a0 = vs + (ve - vs) * c / C, a1 = vs + (ve - vs) * (c + 1) / C;
  // a group of 16 lanes per vertex: 4 vertices in flight per wave
  const int q = threadIdx.x & 15, grp = threadIdx.x >> 4, ng = blockDim.x >> 4;
  for (int64_t k = a0 + grp; k < a1; k += ng) {
    const int32_t v = vperm[k];
    int32_t *row = T + (lab[v] - r0) * B;
    for (int32_t j = optr[v] + q; j < optr[v + 1]; j += 16) atomicAdd(&row[lab[oadj[j]]], ow[j]);
  }
  __syncthreads();
  int32_t *out = partial + c * (int64_t)B * B + (int64_t)r0 * B;
  for (int32_t i = threadIdx.x; i < nt; i += blockDim.x) out[i] = T[i];
}

// Rows of D (blocks 0..B-1) and of its transpose DT (blocks B..2B-1) in one launch: the
// nonzero count of row x into cnt[x] and its sum into deg (d_out for D, d_in for DT)
__global__ __launch_bounds__(256) void k_dense_rows2(const int32_t *__restrict__ D,
                                                     const int32_t *__restrict__ DT, int32_t B,
                                                     uint32_t *__restrict__ cnt,
                                                     int32_t *__restrict__ dout,
                                                     int32_t *__restrict__ din) {
  const int32_t x = (int32_t)(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (x >= 2 * B) return;
  const int lane = threadIdx.x & 63;
  const bool tr = x >= B;
  const int32_t r = tr ? x - B : x;
  const int32_t *row = (tr ? DT : D) + (int64_t)r * B;
  uint32_t c = 0;
  int32_t s = 0;
  for (int32_t base = 0; base < B; base += 256) {
    int32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int32_t t = base + 64 * k + lane;
      v[k] = t < B ? row[t] : 0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      c += v[k] != 0;
      s += v[k];
    }
  }
  for (int off = 32; off >= 1; off >>= 1) {
    c += __shfl_down(c, off, 64);
    s += __shfl_down(s, off, 64);
  }
  if (lane == 0) {
    cnt[x] = c;
    (tr ? din : dout)[r] = s;
  }
}

// The nonzero entries of D's rows (CSR, with erow) and of DT's rows (CSC) at the scanned
// offsets off[0..2B] (columns from off[B]), the row and column pointers, and the log2
// block degrees
__global__ __launch_bounds__(256) void k_dense_write_rows2(
    const int32_t *__restrict__ D, const int32_t *__restrict__ DT, int32_t B,
    const uint32_t *__restrict__ off, int32_t *__restrict__ rowptr, int32_t *__restrict__ col,
    int32_t *__restrict__ val, int32_t *__restrict__ erow, int32_t *__restrict__ colptr,
    int32_t *__restrict__ rowc, int32_t *__restrict__ cval, const int32_t *__restrict__ dout,
    const int32_t *__restrict__ din, double *__restrict__ ldout, double *__restrict__ ldin) {
  const int32_t x = (int32_t)(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (x >= 2 * B) return;
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const bool tr = x >= B;
  const int32_t r = tr ? x - B : x;
  const uint32_t nnz = off[B];
  int32_t run = (int32_t)(tr ? off[x] - nnz : off[x]);
  if (lane == 0) {
    if (tr) {
      colptr[r] = run;
      if (r == B - 1) colptr[B] = (int32_t)(off[2 * B] - nnz);
      ldin[r] = clog2((double)din[r]);
    } else {
      rowptr[r] = run;
      if (r == B - 1) rowptr[B] = (int32_t)nnz;
      ldout[r] = clog2((double)dout[r]);
    }
  }
  const int32_t *row = (tr ? DT : D) + (int64_t)r * B;
  int32_t *ok = tr ? rowc : col, *ov = tr ? cval : val;
  for (int32_t base = 0; base < B; base += 256) {
    int32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int32_t t = base + 64 * k + lane;
      v[k] = t < B ? row[t] : 0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t m = __ballot(v[k] != 0);
      if (v[k] != 0) {
        const int32_t pos = run + (int32_t)__popcll(m & lt);
        ok[pos] = base + 64 * k + lane;
        ov[pos] = v[k];
        if (!tr) erow[pos] = r;
      }
      run += (int32_t)__popcll(m);
    }
  }
}

// DT = D^T through 32x32 LDS tiles (coalesced reads and writes)
__global__ __launch_bounds__(256) void k_transpose(const int32_t *D, int32_t B, int32_t *DT) {
  __shared__ int32_t tile[32][33];
  const int32_t bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int32_t r = by + k, c = bx + tx;
    if (r < B && c < B) tile[k][tx] = D[(int64_t)r * B + c];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int32_t r = bx + k, c = by + tx;
    if (r < B && c < B) DT[(int64_t)r * B + c] = tile[tx][k];
  }
}

// -------------------------------------------------------------- segmented path
// Segment classes by entry count n: registers in groups of 16 or 32 lanes or one wave (one
// entry per lane, n <= 64), one wave with 2, 4 or 8 entries per lane (n <= 512).  Larger segments: when B fits one column range
// (B <= SEG_RC) a workgroup counts the segment's weights per key in LDS and writes the
// nonzero keys in order over the segment's start (one pass); otherwise one workgroup's LDS
// sort takes n <= 4096 and larger segments are counted per (segment, column range) in two
// passes (counts, then writes at the final position).
enum { CL_A16 = 0, CL_A32 = 1, CL_A64 = 2, CL_W128 = 3, CL_W256 = 4, CL_W512 = 5, CL_G1K = 6,
       CL_G2K = 7, CL_G4K = 8, CL_C = 9 };
// SEG_RC counters of a column range plus the workgroup's reduction words stay under 64 KB
constexpr int SEG_W = 512, SEG_G = 4096, SEG_RC = 16000;
constexpr unsigned SEG_GRID = 1024;  // workgroups of the persistent segment kernels
constexpr int SEG_PRIV_B = 4096;     // up to this B, reservations are counted in LDS first
constexpr int SEG_VT = 2048;         // vertices per workgroup of the LDS-counted reservation
// reg_max: largest segment sorted in register groups (16, 32 or 64); count_min: with one
// column range, segments longer than this (512..4096) are counted in LDS, shorter ones sorted
GSAP_HD int seg_class(uint32_t n, bool one_range, uint32_t reg_max = 64,
                      uint32_t count_min = SEG_W) {
  if (n <= reg_max) return n <= 16 ? CL_A16 : n <= 32 ? CL_A32 : CL_A64;
  if (n <= 128) return CL_W128;
  if (n <= 256) return CL_W256;
  if (n <= (uint32_t)SEG_W) return CL_W512;
  if (one_range && n > count_min) return CL_C;
  if (n <= 1024) return CL_G1K;
  if (n <= 2048) return CL_G2K;
  return n <= (uint32_t)SEG_G ? CL_G4K : CL_C;
}

__global__ void k_seg_zero(uint32_t *off, uint32_t *u, uint32_t *ncl, int32_t *dout,
                           int32_t *din, int32_t B) {
  const int64_t n2 = 2 * (int64_t)B;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n2;
       i += (int64_t)gridDim.x * blockDim.x) {
    off[i] = 0;
    if (i < n2) u[i] = 0;
    if (i < B) { dout[i] = 0; din[i] = 0; }
  }
  if (blockIdx.x == 0 && threadIdx.x < SEG_NCL) ncl[threadIdx.x] = 0;
}

// Segment sizes, each vertex's slice start inside its row segment (vrel[v]) and column
// segment (vrel[N + v]), and the weighted block degrees.  PRIV: a workgroup counts its
// SEG_VT vertices per block in LDS and reserves its share of each segment (and adds its
// degree sums) with one global atomic per block, so that blocks of many vertices do not
// serialize on one address.
template <bool PRIV>
__global__ __launch_bounds__(256) void k_seg_reserve(LabView lab, const int32_t *optr,
                                                     const int32_t *iptr, const int32_t *kout,
                                                     const int32_t *kin, int64_t N, int32_t B,
                                                     uint32_t *cnt, int32_t *dout,
                                                     int32_t *din, uint32_t *vrel) {
  if constexpr (!PRIV) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
         v += (int64_t)gridDim.x * blockDim.x) {
      const int32_t r = lab[v];
      const int32_t no = optr[v + 1] - optr[v], ni = iptr[v + 1] - iptr[v];
      vrel[v] = no ? atomicAdd(&cnt[r], (uint32_t)no) : 0u;
      vrel[N + v] = ni ? atomicAdd(&cnt[B + r], (uint32_t)ni) : 0u;
      if (kout) {  // (unit weights: the degrees are the segment sizes, see k_logs_seg)
        if (kout[v]) atomicAdd(&dout[r], kout[v]);
        if (kin[v]) atomicAdd(&din[r], kin[v]);
      }
    }
  } else {
    extern __shared__ uint32_t sh[];
    uint32_t *cr = sh, *cc = sh + B, *so = sh + 2 * B, *si = sh + 3 * B;
    for (int32_t i = threadIdx.x; i < 4 * B; i += 256) sh[i] = 0;
    __syncthreads();
    constexpr int PER = SEG_VT / 256;
    const int64_t v0 = (int64_t)blockIdx.x * SEG_VT + threadIdx.x;
    int32_t rr[PER];
    uint32_t ro[PER], ri[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int64_t v = v0 + k * 256;
      rr[k] = -1;
      ro[k] = ri[k] = 0;
      if (v < N) {
        const int32_t r = lab[v];
        const int32_t no = optr[v + 1] - optr[v], ni = iptr[v + 1] - iptr[v];
        rr[k] = r;
        if (no) ro[k] = atomicAdd(&cr[r], (uint32_t)no);
        if (ni) ri[k] = atomicAdd(&cc[r], (uint32_t)ni);
        if (kout[v]) atomicAdd(&so[r], (uint32_t)kout[v]);
        if (kin[v]) atomicAdd(&si[r], (uint32_t)kin[v]);
      }
    }
    __syncthreads();
    for (int32_t b = threadIdx.x; b < B; b += 256) {
      if (cr[b]) cr[b] = atomicAdd(&cnt[b], cr[b]);
      if (cc[b]) cc[b] = atomicAdd(&cnt[B + b], cc[b]);
      if (so[b]) atomicAdd(&dout[b], (int32_t)so[b]);
      if (si[b]) atomicAdd(&din[b], (int32_t)si[b]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int64_t v = v0 + k * 256;
      if (rr[k] >= 0) {
        vrel[v] = cr[rr[k]] + ro[k];
        vrel[N + v] = cc[rr[k]] + ri[k];
      }
    }
  }
}

// Edge-parallel: out-edge e goes to its source's slice of the source block's row segment,
// keyed by the target's block; in-edge f to its target's slice of the target block's column
// segment, keyed by the source's block.  With wb > 0 an entry is one word, key << wb | weight
// (ew untouched).  ow = iw = nullptr: unit weights.
__global__ __launch_bounds__(256) void k_seg_fill(LabView lab, const int32_t *esrc,
                                                  const int32_t *optr, const int32_t *oadj,
                                                  const int32_t *ow, const int32_t *idst,
                                                  const int32_t *iptr, const int32_t *iadj,
                                                  const int32_t *iw, int64_t N, int64_t E,
                                                  int32_t B, const uint32_t *off,
                                                  const uint32_t *vrel, uint32_t *ekey,
                                                  uint32_t *ew, int wb) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 2 * E;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (e < E) {
      const int32_t v = esrc[e];
      const uint32_t x = (uint32_t)lab[v];
      const uint32_t pos = off[x] + vrel[v] + (uint32_t)(e - optr[v]);
      const uint32_t k = (uint32_t)lab[oadj[e]], w = ow ? (uint32_t)ow[e] : 1u;
      if (wb) {
        ekey[pos] = k << wb | w;
      } else {
        ekey[pos] = k;
        ew[pos] = w;
      }
    } else {
      const int64_t f = e - E;
      const int32_t v = idst[f];
      const uint32_t x = (uint32_t)(B + lab[v]);
      const uint32_t pos = off[x] + vrel[N + v] + (uint32_t)(f - iptr[v]);
      const uint32_t k = (uint32_t)lab[iadj[f]], w = iw ? (uint32_t)iw[f] : 1u;
      if (wb) {
        ekey[pos] = k << wb | w;
      } else {
        ekey[pos] = k;
        ew[pos] = w;
      }
    }
  }
}

// Nonempty segments into per-class lists: per-wave ballots, per-workgroup LDS counts, one
// global atomic per class and workgroup round (1024 segments).  first (when not null): the
// segment holding position 64 c of the segment array, for every chunk c (the thread of a
// segment writes the chunks that start inside it), read by the compaction.
__global__ __launch_bounds__(256) void k_seg_classify(const uint32_t *off, int64_t nseg,
                                                      int one_range, uint32_t reg_max,
                                                      uint32_t count_min, uint32_t *lists,
                                                      int64_t lcap, uint32_t *ncl,
                                                      uint32_t *first) {
  __shared__ uint32_t wcnt[SEG_NCL], wbase[SEG_NCL];
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int64_t base = (int64_t)blockIdx.x * 1024; base < nseg; base += (int64_t)gridDim.x * 1024) {
    if (threadIdx.x < SEG_NCL) wcnt[threadIdx.x] = 0;
    __syncthreads();
    int c[4];
    uint32_t woff[4][SEG_NCL];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t x = base + j * 256 + threadIdx.x;
      c[j] = -1;
      if (x < nseg) {
        const uint32_t s0 = off[x], s1 = off[x + 1], n = s1 - s0;
        if (n) c[j] = seg_class(n, one_range != 0, reg_max, count_min);
        if (first)
          for (uint32_t ch = (s0 + 63) >> 6; (ch << 6) < s1; ++ch) first[ch] = (uint32_t)x;
      }
#pragma unroll
      for (int k = 0; k < SEG_NCL; ++k) {
        const uint64_t m = __ballot(c[j] == k);
        uint32_t o = 0;
        if (m && lane == 0) o = atomicAdd(&wcnt[k], (uint32_t)__popcll(m));
        woff[j][k] = __shfl(o, 0, 64) + (uint32_t)__popcll(m & lt);
      }
    }
    __syncthreads();
    if (threadIdx.x < SEG_NCL)
      wbase[threadIdx.x] = wcnt[threadIdx.x] ? atomicAdd(&ncl[threadIdx.x], wcnt[threadIdx.x]) : 0u;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t x = base + j * 256 + threadIdx.x;
#pragma unroll
      for (int k = 0; k < SEG_NCL; ++k)
        if (c[j] == k) lists[k * lcap + wbase[k] + woff[j][k]] = (uint32_t)x;
    }
    __syncthreads();
  }
}

// Sort a segment of n <= G entries held one per lane of a G-lane group (bitonic network on
// shuffles), then reduce runs of equal keys: run totals by a segmented suffix sum, heads
// compacted by ballot.  Writes the runs in place at the segment start and their count.
// PK: packed entries (key << wb | weight) sorted as single words, split after the network.
template <int G, bool PK>
GSAP_D void seg_sort_group(uint32_t *ekey, uint32_t *ew, uint32_t s0, int n, uint32_t *uout,
                           int wb) {
  const int q = threadIdx.x & (G - 1);
  const int gbase = (threadIdx.x & 63) & ~(G - 1);
  const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
  uint32_t key = q < n ? ekey[s0 + q] : 0xFFFFFFFFu;
  uint32_t val = 0u;
  if constexpr (!PK) val = q < n ? ew[s0 + q] : 0u;
#pragma unroll
  for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint32_t pk = __shfl_xor(key, j, G);
      const uint32_t mmin = ((q & k) == 0) == ((q & j) == 0) ? ~0u : 0u;
      const uint32_t nk = (min(pk, key) & mmin) | (max(pk, key) & ~mmin);
      if constexpr (!PK) {
        const uint32_t pv = __shfl_xor(val, j, G);
        val = nk != key ? pv : val;
      }
      key = nk;
    }
  }
  if constexpr (PK) {
    // the padding words sort after every entry: lanes q >= n hold padding only
    val = q < n ? key & ((1u << wb) - 1u) : 0u;
    key >>= wb;
  }
  uint32_t acc = val;
#pragma unroll
  for (int o = 1; o < G; o <<= 1) {
    const uint32_t y = __shfl_down(acc, o, G);
    const uint32_t ky = __shfl_down(key, o, G);
    if (q + o < G && ky == key) acc += y;
  }
  const uint32_t kp = __shfl_up(key, 1, G);
  const bool head = q < n && (q == 0 || kp != key);
  const uint64_t hm = (__ballot(head) >> gbase) & gmask;
  if (head) {
    const uint32_t pos = (uint32_t)__popcll(hm & ((1ull << q) - 1ull));
    ekey[s0 + pos] = key;
    ew[s0 + pos] = acc;
  }
  if (q == 0) *uout = (uint32_t)__popcll(hm);
}

// Wave tasks: one segment of <= 64 entries, two of <= 32 or four of <= 16.
template <bool PK>
__global__ __launch_bounds__(256) void k_seg_sort_reg(const uint32_t *off, uint32_t *ekey,
                                                      uint32_t *ew, const uint32_t *lists,
                                                      int64_t lcap, const uint32_t *ncl,
                                                      uint32_t *u, int wb) {
  const int64_t n64 = ncl[CL_A64], n32 = ncl[CL_A32], n16 = ncl[CL_A16];
  const int64_t t32 = (n32 + 1) / 2, t16 = (n16 + 3) / 4;
  const int64_t tasks = n64 + t32 + t16;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int lane = threadIdx.x & 63;
  for (int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < tasks; t += nw) {
    if (t < n64) {
      const uint32_t x = lists[CL_A64 * lcap + t];
      seg_sort_group<64, PK>(ekey, ew, off[x], (int)(off[x + 1] - off[x]), &u[x], wb);
    } else if (t < n64 + t32) {
      const int64_t k = (t - n64) * 2 + (lane >> 5);
      if (k < n32) {
        const uint32_t x = lists[CL_A32 * lcap + k];
        seg_sort_group<32, PK>(ekey, ew, off[x], (int)(off[x + 1] - off[x]), &u[x], wb);
      }
    } else {
      const int64_t k = (t - n64 - t32) * 4 + (lane >> 4);
      if (k < n16) {
        const uint32_t x = lists[CL_A16 * lcap + k];
        seg_sort_group<16, PK>(ekey, ew, off[x], (int)(off[x + 1] - off[x]), &u[x], wb);
      }
    }
  }
}

// inclusive scan of one value per thread over a team of T threads (one wave, or a 256-thread
// workgroup using sc4)
template <int T>
GSAP_D uint32_t team_incl_scan(uint32_t x, uint32_t *sc4, uint32_t &total) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if constexpr (T == 64) {
    total = __shfl(x, 63, 64);
    return x;
  } else {
    const int w = threadIdx.x >> 6;
    if (lane == 63) sc4[w] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < T / 64; ++i) {
      const uint32_t c = sc4[i];
      if (i < w) pre += c;
      tot += c;
    }
    __syncthreads();
    total = tot;
    return pre + x;
  }
}

// Sort a segment of n <= T*R entries held in registers, R per thread in blocked order
// (element g = thread * R + i), with a bitonic network: steps with j < R compare registers of
// one thread, R <= j < 64R exchange with lane (lane ^ j/R) by shuffles, larger j (T = 256)
// go through LDS (xk, xv: T*R words each).  Then runs of equal keys are reduced: run heads
// record the exclusive prefix of the weights in xv[run], run tails write key and total to
// the segment's start.  Padding keys are 0xFFFFFFFF (above every block id, and above every
// packed entry).  PK: packed entries (key << wb | weight) sorted as single words, split after
// the network.
template <int T, int R, bool PK>
GSAP_D void team_sort_regs(uint32_t *ekey, uint32_t *ew, uint32_t s0, int n, uint32_t *uout,
                           uint32_t *xk, uint32_t *xv, uint32_t *sc4, int wb) {
  const int tid = T == 64 ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
  constexpr int P = T * R;
  uint32_t key[R], val[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int g = tid * R + i;
    key[i] = g < n ? ekey[s0 + g] : 0xFFFFFFFFu;
    val[i] = 0u;
    if constexpr (!PK) val[i] = g < n ? ew[s0 + g] : 0u;
  }
#pragma unroll
  for (int k = 2; k <= P; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j < R) {
        // branch-free compare-exchange of registers i < l = i ^ j
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const int l = i ^ j;
          if (l > i) {
            const bool asc = ((tid * R + i) & k) == 0;
            const uint32_t a = key[i], b = key[l];
            const bool sw = (a > b) == asc;
            key[i] = sw ? b : a;
            key[l] = sw ? a : b;
            if constexpr (!PK) {
              const uint32_t va = val[i], vb = val[l];
              val[i] = sw ? vb : va;
              val[l] = sw ? va : vb;
            }
          }
        }
      } else if (j < 64 * R) {
        // all of the step's shuffles first, then branch-free selects; which side keeps the
        // minimum depends on lane bits only (j, k >= R)
        const int lm = j / R;
        uint32_t pk[R], pv[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
          pk[i] = __shfl_xor(key[i], lm, 64);
          if constexpr (!PK) pv[i] = __shfl_xor(val[i], lm, 64);
        }
        const int g0 = tid * R;
        const uint32_t mmin = ((g0 & j) == 0) == ((g0 & k) == 0) ? ~0u : 0u;
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const uint32_t nk = (min(pk[i], key[i]) & mmin) | (max(pk[i], key[i]) & ~mmin);
          if constexpr (!PK) val[i] = nk != key[i] ? pv[i] : val[i];
          key[i] = nk;
        }
      } else {
#pragma unroll
        for (int i = 0; i < R; ++i) {
          xk[tid * R + i] = key[i];
          if constexpr (!PK) xv[tid * R + i] = val[i];
        }
        __syncthreads();
        const int g0 = tid * R;
        const uint32_t mmin = ((g0 & j) == 0) == ((g0 & k) == 0) ? ~0u : 0u;
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const uint32_t pk = xk[(g0 + i) ^ j];
          const uint32_t nk = (min(pk, key[i]) & mmin) | (max(pk, key[i]) & ~mmin);
          if constexpr (!PK) val[i] = nk != key[i] ? xv[(g0 + i) ^ j] : val[i];
          key[i] = nk;
        }
        __syncthreads();
      }
    }
  }
  if constexpr (PK) {
    // the padding words sort after every entry: positions g >= n hold padding only
#pragma unroll
    for (int i = 0; i < R; ++i) {
      val[i] = tid * R + i < n ? key[i] & ((1u << wb) - 1u) : 0u;
      key[i] >>= wb;
    }
  }
  // neighbours of the thread's first and last element
  uint32_t kprev, knext;
  if constexpr (T == 64) {
    kprev = __shfl_up(key[R - 1], 1, 64);
    knext = __shfl_down(key[0], 1, 64);
  } else {
    xk[tid] = key[R - 1];
    xk[T + tid] = key[0];
    __syncthreads();
    kprev = tid > 0 ? xk[tid - 1] : 0u;
    knext = tid + 1 < T ? xk[T + tid + 1] : 0xFFFFFFFFu;
    __syncthreads();
  }
  uint32_t hsum = 0, wsum = 0;
  bool head[R], tail[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int g = tid * R + i;
    const uint32_t kp = i > 0 ? key[i - 1] : kprev;
    const uint32_t kn = i + 1 < R ? key[i + 1] : knext;
    head[i] = g < n && (g == 0 || kp != key[i]);
    tail[i] = g < n && (g == n - 1 || kn != key[i]);
    hsum += head[i] ? 1u : 0u;
    wsum += val[i];
  }
  uint32_t th, tw;
  const uint32_t hi = team_incl_scan<T>(hsum, sc4, th);
  const uint32_t wi = team_incl_scan<T>(wsum, sc4, tw);
  uint32_t run = hi - hsum, ex = wi - wsum;  // heads and weight before this thread
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (head[i]) xv[run] = ex;
    run += head[i] ? 1u : 0u;
    ex += val[i];
  }
  if constexpr (T == 64) wave_lds_sync();
  else __syncthreads();
  run = hi - hsum;
  ex = wi - wsum;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    run += head[i] ? 1u : 0u;
    if (tail[i]) {
      ekey[s0 + run - 1] = key[i];
      ew[s0 + run - 1] = (ex + val[i]) - xv[run - 1];
    }
    ex += val[i];
  }
  if (tid == 0) *uout = th;
  if constexpr (T == 64) wave_lds_sync();
  else __syncthreads();
}

// one wave per segment of 65..512 entries: 2, 4 or 8 entries per lane (networks of 128,
// 256 or 512), 2 KB of LDS per wave
template <bool PK>
__global__ __launch_bounds__(256) void k_seg_sort_w(const uint32_t *off, uint32_t *ekey,
                                                    uint32_t *ew, const uint32_t *lists,
                                                    int64_t lcap, const uint32_t *ncl,
                                                    uint32_t *u, int wb) {
  __shared__ uint32_t lds[4][SEG_W];
  const int w = threadIdx.x >> 6;
  const int64_t n512 = ncl[CL_W512], n256 = ncl[CL_W256], n128 = ncl[CL_W128];
  const int64_t tasks = n512 + n256 + n128;
  for (int64_t t = (int64_t)blockIdx.x * 4 + w; t < tasks; t += (int64_t)gridDim.x * 4) {
    if (t < n512) {
      const uint32_t x = lists[CL_W512 * lcap + t];
      team_sort_regs<64, 8, PK>(ekey, ew, off[x], (int)(off[x + 1] - off[x]), &u[x], nullptr,
                                lds[w], nullptr, wb);
    } else if (t < n512 + n256) {
      const uint32_t x = lists[CL_W256 * lcap + (t - n512)];
      team_sort_regs<64, 4, PK>(ekey, ew, off[x], (int)(off[x + 1] - off[x]), &u[x], nullptr,
                                lds[w], nullptr, wb);
    } else {
      const uint32_t x = lists[CL_W128 * lcap + (t - n512 - n256)];
      team_sort_regs<64, 2, PK>(ekey, ew, off[x], (int)(off[x + 1] - off[x]), &u[x], nullptr,
                                lds[w], nullptr, wb);
    }
  }
}

// one workgroup per segment of 513..4096 entries: 4, 8 or 16 entries per thread (networks
// of 1024, 2048 or 4096), 32 KB of LDS
// One size class per instance (its own register and LDS budget: the 1K and 2K classes run
// at a higher occupancy than the 4K class would allow them).
template <int R, int CL, bool PK>
__global__ __launch_bounds__(256) void k_seg_sort_g(const uint32_t *off, uint32_t *ekey,
                                                    uint32_t *ew, const uint32_t *lists,
                                                    int64_t lcap, const uint32_t *ncl,
                                                    uint32_t *u, int wb) {
  __shared__ uint32_t xk[256 * R], xv[256 * R];
  __shared__ uint32_t sc4[4];
  const int64_t n = ncl[CL];
  for (int64_t t = blockIdx.x; t < n; t += gridDim.x) {
    const uint32_t x = lists[CL * lcap + t];
    team_sort_regs<256, R, PK>(ekey, ew, off[x], (int)(off[x + 1] - off[x]), &u[x], xk, xv, sc4,
                               wb);
  }
}

// Class C: weights of the segment's keys in [c0, c0 + cn) counted in LDS by a workgroup of
// 512 threads (four entries in flight per thread).  Ends with a barrier: every read of the
// segment is complete.  PK: packed entries (key << wb | weight).
template <bool PK>
GSAP_D void seg_range_count(uint32_t *cnt, const uint32_t *ekey, const uint32_t *ew,
                            uint32_t s0, uint32_t s1, uint32_t c0, uint32_t cn, int wb) {
  for (uint32_t i = threadIdx.x; i < cn; i += 512) cnt[i] = 0;
  __syncthreads();
  for (uint64_t e0 = (uint64_t)s0 + threadIdx.x; e0 < s1; e0 += 4 * 512) {
    uint32_t k[4], w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint64_t e = e0 + (uint64_t)j * 512;
      if constexpr (PK) {
        const uint32_t p = e < s1 ? ekey[e] : 0u;
        k[j] = e < s1 ? (p >> wb) - c0 : 0xFFFFFFFFu;
        w[j] = p & ((1u << wb) - 1u);
      } else {
        k[j] = e < s1 ? ekey[e] - c0 : 0xFFFFFFFFu;
        w[j] = e < s1 ? ew[e] : 0u;
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (k[j] < cn) atomicAdd(&cnt[k[j]], w[j]);
  }
  __syncthreads();
}

GSAP_D uint32_t block_sum512(uint32_t x, uint32_t *red8) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
  if (lane == 0) red8[w] = x;
  __syncthreads();
  uint32_t t = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) t += red8[i];
  __syncthreads();
  return t;
}

// Nonzero counters cnt[0..cn) in ascending order to keys c0 + i / weights at out + pos
// (and the matching dense row, and the coarse index entry of every 16th position, when
// given); returns the number written.
GSAP_D uint32_t seg_emit(const uint32_t *cnt, uint32_t cn, uint32_t c0, uint32_t pos,
                         uint32_t *red8, uint32_t *okey, uint32_t *ow_, int32_t *col,
                         int32_t *val, int32_t *erow, int32_t x_row, int32_t *dense_row,
                         int32_t *gidx = nullptr) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t start = pos;
  for (uint32_t base = 0; base < cn; base += 512) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t c = i < cn ? cnt[i] : 0u;
    const uint64_t m = __ballot(c != 0);
    if (lane == 0) red8[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t cj = red8[j];
      if (j < w) wpre += cj;
      tot += cj;
    }
    __syncthreads();
    if (c != 0) {
      const uint32_t p = pos + wpre + (uint32_t)__popcll(m & lt);
      if (okey) {
        okey[p] = c0 + i;
        ow_[p] = c;
      } else {
        col[p] = (int32_t)(c0 + i);
        val[p] = (int32_t)c;
        if (erow) erow[p] = x_row;
        if (gidx && (p & 15u) == 0) gidx[p >> 4] = (int32_t)(c0 + i);
      }
      if (dense_row) dense_row[c0 + i] = (int32_t)c;
    }
    pos += tot;
  }
  return pos - start;
}

// B <= SEG_RC: one workgroup per class-C segment, all B keys in one LDS range; the runs
// overwrite the segment's start (its reads are complete after the count).
template <bool PK>
__global__ __launch_bounds__(512) void k_seg_c1(const uint32_t *off, uint32_t *ekey,
                                                uint32_t *ew, const uint32_t *lists,
                                                int64_t lcap, const uint32_t *ncl, int32_t B,
                                                uint32_t *u, int wb) {
  extern __shared__ uint32_t cnt[];
  __shared__ uint32_t red8[8];
  const int64_t n = ncl[CL_C];
  for (int64_t t = blockIdx.x; t < n; t += gridDim.x) {
    const uint32_t x = lists[CL_C * lcap + t];
    const uint32_t s0 = off[x];
    seg_range_count<PK>(cnt, ekey, ew, s0, off[x + 1], 0, (uint32_t)B, wb);
    const uint32_t m = seg_emit(cnt, (uint32_t)B, 0, s0, red8, ekey, ew, nullptr, nullptr,
                                nullptr, 0, nullptr);
    if (threadIdx.x == 0) u[x] = m;
  }
}

// B > SEG_RC, pass 1: nonzero count of every (class-C segment, column range)
template <bool PK>
__global__ __launch_bounds__(512) void k_seg_c_count(const uint32_t *off, const uint32_t *ekey,
                                                     const uint32_t *ew, const uint32_t *lists,
                                                     int64_t lcap, const uint32_t *ncl,
                                                     int32_t B, int32_t nr, uint32_t *cntc,
                                                     uint32_t *u, int wb) {
  extern __shared__ uint32_t cnt[];
  __shared__ uint32_t red8[8];
  const uint32_t RC = B < SEG_RC ? (uint32_t)B : (uint32_t)SEG_RC;
  const int64_t items = (int64_t)ncl[CL_C] * nr;
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int64_t k = it / nr;
    const uint32_t q = (uint32_t)(it - k * nr);
    const uint32_t x = lists[CL_C * lcap + k];
    const uint32_t c0 = q * RC, cn = min(RC, (uint32_t)B - c0);
    seg_range_count<PK>(cnt, ekey, ew, off[x], off[x + 1], c0, cn, wb);
    uint32_t nz = 0;
    for (uint32_t i = threadIdx.x; i < cn; i += 512) nz += cnt[i] != 0;
    nz = block_sum512(nz, red8);
    if (threadIdx.x == 0) {
      cntc[it] = nz;
      if (nz) atomicAdd(&u[x], nz);
    }
  }
}

// B > SEG_RC, pass 2: recount each (segment, range) and write its nonzeros at the
// segment's final position plus the earlier ranges' counts
template <bool PK>
__global__ __launch_bounds__(512) void k_seg_c_write(const uint32_t *off, const uint32_t *ekey,
                                                     const uint32_t *ew, const uint32_t *lists,
                                                     int64_t lcap, const uint32_t *ncl,
                                                     int32_t B, int32_t nr,
                                                     const uint32_t *cntc, const uint32_t *uoff,
                                                     int32_t *col, int32_t *val, int32_t *erow,
                                                     int32_t *row, int32_t *cval,
                                                     int32_t *dense, int32_t *gcol,
                                                     int32_t *grow, int wb) {
  extern __shared__ uint32_t cnt[];
  __shared__ uint32_t red8[8];
  const uint32_t RC = B < SEG_RC ? (uint32_t)B : (uint32_t)SEG_RC;
  const int64_t items = (int64_t)ncl[CL_C] * nr;
  const uint32_t nnz = uoff[B];
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int64_t k = it / nr;
    const uint32_t q = (uint32_t)(it - k * nr);
    if (cntc[it] == 0) continue;  // uniform across the workgroup
    const uint32_t x = lists[CL_C * lcap + k];
    const bool isrow = x < (uint32_t)B;
    uint32_t pre = 0;
    for (uint32_t qq = threadIdx.x; qq < q; qq += 512) pre += cntc[k * nr + qq];
    pre = block_sum512(pre, red8);
    const uint32_t pos = (isrow ? uoff[x] : uoff[x] - nnz) + pre;
    const uint32_t c0 = q * RC, cn = min(RC, (uint32_t)B - c0);
    seg_range_count<PK>(cnt, ekey, ew, off[x], off[x + 1], c0, cn, wb);
    if (isrow)
      seg_emit(cnt, cn, c0, pos, red8, nullptr, nullptr, col, val, erow, (int32_t)x,
               dense ? dense + (int64_t)x * B : nullptr, gcol);
    else
      seg_emit(cnt, cn, c0, pos, red8, nullptr, nullptr, row, cval, nullptr, 0, nullptr, grow);
  }
}

// Row/column pointers from the scanned run counts, then entry-parallel: the runs of every
// segment (except two-pass class C, written at their place already) copied from the
// segment's start to its final position, four entries per thread.  An entry's segment is
// found from its chunk's first segment by a short forward walk over the segment offsets.
__global__ __launch_bounds__(256) void k_seg_compact(const uint32_t *off, const uint32_t *ekey,
                                                     const uint32_t *ew, const uint32_t *first,
                                                     const uint32_t *uoff, int32_t B,
                                                     int one_range, int64_t n_ent,
                                                     int32_t *rowptr, int32_t *col,
                                                     int32_t *val, int32_t *erow,
                                                     int32_t *colptr, int32_t *row,
                                                     int32_t *cval, int32_t *dense,
                                                     int32_t *gcol, int32_t *grow) {
  const uint32_t nnz = uoff[B];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t x = tid; x <= 2 * (int64_t)B; x += stride) {
    if (x < B) rowptr[x] = (int32_t)uoff[x];
    else if (x < 2 * (int64_t)B) colptr[x - B] = (int32_t)(uoff[x] - nnz);
    else { rowptr[B] = (int32_t)nnz; colptr[B] = (int32_t)nnz; }
  }
  for (int64_t e0 = tid; e0 < n_ent; e0 += 4 * stride) {
    uint32_t x[4], key[4], w[4], s0[4];
    bool ok[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t e = e0 + j * stride;
      ok[j] = e < n_ent;
      if (!ok[j]) continue;
      x[j] = first[e >> 6];
      key[j] = ekey[e];
      w[j] = ew[e];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!ok[j]) continue;
      const uint32_t e = (uint32_t)(e0 + j * stride);
      while (off[x[j] + 1] <= e) ++x[j];
      s0[j] = off[x[j]];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!ok[j]) continue;
      const uint32_t e = (uint32_t)(e0 + j * stride);
      const uint32_t n = off[x[j] + 1] - s0[j];
      const uint32_t k = e - s0[j];
      const uint32_t cnt = uoff[x[j] + 1] - uoff[x[j]];
      if (k >= cnt || (!one_range && seg_class(n, false) == CL_C)) continue;
      const uint32_t q = uoff[x[j]] - (x[j] < (uint32_t)B ? 0u : nnz) + k;
      if (x[j] < (uint32_t)B) {
        col[q] = (int32_t)key[j];
        val[q] = (int32_t)w[j];
        erow[q] = (int32_t)x[j];
        if (dense) dense[(int64_t)x[j] * B + key[j]] = (int32_t)w[j];
        if (gcol && (q & 15u) == 0) gcol[q >> 4] = (int32_t)key[j];
      } else {
        row[q] = (int32_t)key[j];
        cval[q] = (int32_t)w[j];
        if (grow && (q & 15u) == 0) grow[q >> 4] = (int32_t)key[j];
      }
    }
  }
}

static unsigned grid_for(int64_t n, int threads, unsigned cap = 65535u * 4u) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > (int64_t)cap) g = cap;
  return (unsigned)g;
}

int64_t seg_cntc_words(int64_t N, int64_t E) {
  const int64_t nr = (N + SEG_RC - 1) / SEG_RC;
  return (2 * E / (SEG_G + 1) + 1) * nr;
}

// Bits of a packed segment entry's weight (key << wb | weight) when every entry and the
// padding key fit: wmax < 2^wb and B << wb < 2^32 (then 0xFFFFFFFF lies above every packed
// entry and, shifted, above every block id); 0 otherwise (keys and weights in two words).
static int seg_pack_bits(int32_t B, int32_t wmax) {
  int b = 1;
  while (b < 31 && ((uint32_t)wmax >> b) != 0) ++b;
  return ((uint64_t)B << b) < (1ull << 32) ? b : 0;
}

template <bool PK>
static void seg_path_t(const GraphDev &g, LabView lab, int32_t B, BMBuf &bm,
                       RebuildScratch &rs, bool dense, int wb, hipStream_t st) {
  const int64_t N = g.N, E = g.E, nseg = 2 * (int64_t)B;
  const int T = 256;
  uint32_t *ekey = (uint32_t *)rs.k0, *ew = (uint32_t *)rs.k1;
  const bool one = B <= SEG_RC;
  const int32_t nr = (int32_t)((B + SEG_RC - 1) / SEG_RC);
  const uint32_t RC = one ? (uint32_t)B : (uint32_t)SEG_RC;
  int32_t *D = dense ? bm.dense : nullptr;
  // without the dense copy, the coarse lookup index is written with the entries
  int32_t *gcol = dense ? nullptr : bm.gcol, *grow = dense ? nullptr : bm.grow;
  // unit weights are not read (degrees = segment sizes, entry weights = 1)
  const bool unit = g.wmax == 1;
  k_seg_zero<<<grid_for(nseg + 1, T), T, 0, st>>>(rs.soff, rs.su, rs.ncl, bm.dout, bm.din, B);
  if (B <= SEG_PRIV_B)
    k_seg_reserve<true><<<(unsigned)((N + SEG_VT - 1) / SEG_VT), 256, (size_t)16 * B, st>>>(
        lab, g.optr, g.iptr, g.kout, g.kin, N, B, rs.soff, bm.dout, bm.din, rs.vrel);
  else
    k_seg_reserve<false><<<grid_for(N, T), T, 0, st>>>(lab, g.optr, g.iptr,
                                                       unit ? nullptr : g.kout,
                                                       unit ? nullptr : g.kin, N, B, rs.soff,
                                                       bm.dout, bm.din, rs.vrel);
  scan_exclusive(rs.soff, rs.soff, nseg, rs.soff + nseg, rs.tmp, st);
  k_seg_fill<<<grid_for(2 * E, T), T, 0, st>>>(lab, g.esrc, g.optr, g.oadj, unit ? nullptr : g.ow,
                                               g.idst, g.iptr, g.iadj, unit ? nullptr : g.iw, N, E,
                                               B, rs.soff, rs.vrel, ekey, ew, wb);
  const uint32_t reg_max = (uint32_t)rs.seg_reg_max, count_min = (uint32_t)rs.seg_count_min;
  k_seg_classify<<<grid_for(nseg, T, SEG_GRID), T, 0, st>>>(rs.soff, nseg, one ? 1 : 0, reg_max,
                                                            count_min, rs.lists, rs.lcap, rs.ncl,
                                                            rs.chunk_seg);
  if (one) {
    k_seg_c1<PK><<<SEG_GRID, 512, (size_t)RC * 4, st>>>(rs.soff, ekey, ew, rs.lists, rs.lcap,
                                                        rs.ncl, B, rs.su, wb);
    // segments of 513..count_min entries sorted instead of counted
    if (count_min > (uint32_t)SEG_W) {
      k_seg_sort_g<16, CL_G4K, PK><<<SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists,
                                                             rs.lcap, rs.ncl, rs.su, wb);
      k_seg_sort_g<8, CL_G2K, PK><<<SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists, rs.lcap,
                                                            rs.ncl, rs.su, wb);
      k_seg_sort_g<4, CL_G1K, PK><<<2 * SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists,
                                                                rs.lcap, rs.ncl, rs.su, wb);
    }
  } else {
    k_seg_sort_g<16, CL_G4K, PK><<<SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists, rs.lcap,
                                                           rs.ncl, rs.su, wb);
    k_seg_sort_g<8, CL_G2K, PK><<<SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists, rs.lcap,
                                                          rs.ncl, rs.su, wb);
    k_seg_sort_g<4, CL_G1K, PK><<<2 * SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists,
                                                              rs.lcap, rs.ncl, rs.su, wb);
    k_seg_c_count<PK><<<SEG_GRID / 2, 512, (size_t)RC * 4, st>>>(
        rs.soff, ekey, ew, rs.lists, rs.lcap, rs.ncl, B, nr, rs.cntc, rs.su, wb);
  }
  k_seg_sort_w<PK><<<SEG_GRID * 3 / 2, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists, rs.lcap,
                                                   rs.ncl, rs.su, wb);
  k_seg_sort_reg<PK><<<SEG_GRID * 2, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists, rs.lcap, rs.ncl,
                                                   rs.su, wb);
  scan_exclusive(rs.su, rs.suoff, nseg, rs.suoff + nseg, rs.tmp, st);
  k_seg_compact<<<grid_for((2 * E + 3) / 4, T), T, 0, st>>>(
      rs.soff, ekey, ew, rs.chunk_seg, rs.suoff, B, one ? 1 : 0, 2 * E, bm.rowptr, bm.col,
      bm.val, bm.erow, bm.colptr, bm.row, bm.cval, D, gcol, grow);
  if (!one)
    k_seg_c_write<PK><<<SEG_GRID / 2, 512, (size_t)RC * 4, st>>>(
        rs.soff, ekey, ew, rs.lists, rs.lcap, rs.ncl, B, nr, rs.cntc, rs.suoff, bm.col, bm.val,
        bm.erow, bm.row, bm.cval, D, gcol, grow, wb);
  if (unit)
    k_logs_seg<<<grid_for(B, 256), 256, 0, st>>>(rs.soff, B, bm.dout, bm.din, bm.ldout, bm.ldin);
  else
    k_logs<<<grid_for(B, 256), 256, 0, st>>>(bm.dout, bm.din, B, bm.ldout, bm.ldin);
}

static void seg_path(const GraphDev &g, LabView lab, int32_t B, BMBuf &bm, RebuildScratch &rs,
                     bool dense, hipStream_t st) {
  const int wb = seg_pack_bits(B, g.wmax);
  if (wb) seg_path_t<true>(g, lab, B, bm, rs, dense, wb, st);
  else seg_path_t<false>(g, lab, B, bm, rs, dense, 0, st);
}

// 128 < B <= tile_max_B (dense): vertices grouped by block; workgroup (row tile, chunk) counts
// the out-edges of its rows' vertices in LDS and writes a partial plane; the planes are
// summed into D; CSR from D's rows, CSC from the rows of D's transpose.
// CSR, CSC, degrees and their logs from the dense matrix D = bm.dense (B x B); the
// transpose goes through the partial-plane buffer (B x B words)
static void dense_to_csr(int32_t B, BMBuf &bm, RebuildScratch &rs, hipStream_t st) {
  const int32_t *D = bm.dense;
  int32_t *DT = rs.partials;
  if (bm.dt_B != B) {  // else the incremental update kept the transpose current
    const unsigned tb = (unsigned)((B + 31) / 32);
    k_transpose<<<dim3(tb, tb), 256, 0, st>>>(D, B, DT);
    bm.dt_B = B;
  }
  // rows of D then rows of DT: counts, one scan over both, entries and pointers
  uint32_t *cnt = rs.su, *off = rs.suoff;
  const unsigned rb = (unsigned)((2 * B + 3) / 4);
  k_dense_rows2<<<rb, 256, 0, st>>>(D, DT, B, cnt, bm.dout, bm.din);
  scan_exclusive(cnt, off, 2 * (int64_t)B, off + 2 * (int64_t)B, rs.tmp, st);
  k_dense_write_rows2<<<rb, 256, 0, st>>>(D, DT, B, off, bm.rowptr, bm.col, bm.val, bm.erow,
                                          bm.colptr, bm.row, bm.cval, bm.dout, bm.din,
                                          bm.ldout, bm.ldin);
}

static void tile_path(const GraphDev &g, LabView lab, int32_t B, BMBuf &bm,
                      RebuildScratch &rs, hipStream_t st) {
  const int64_t N = g.N;
  const int64_t BB = (int64_t)B * B;
  const int T = 256;
  int32_t *D = bm.dense;
  uint32_t *vhist = rs.lists, *cnt = rs.su;
  int32_t *vperm = (int32_t *)rs.vrel;
  const int32_t R = 16384 / B;
  const int32_t ntiles = (B + R - 1) / R;
  const int32_t W = (int32_t)((N + VT - 1) / VT);
  k_label_hist<<<W, 256, (size_t)B * 4, st>>>(lab, N, B, vhist, W);
  scan_exclusive(vhist, vhist, (int64_t)B * W, nullptr, rs.tmp, st);
  k_label_scatter<<<W, 256, (size_t)B * 4, st>>>(lab, N, B, vhist, W, vperm);
  // about 2048 workgroups, within a budget of partial counters that grows with E (a
  // workgroup's chunk stays near E / 1024 edges at most sizes)
  int32_t C = 2048 / ntiles;
  const int64_t budget = std::min<int64_t>(std::max<int64_t>((int64_t)4 << 20, g.E / 2), rs.partials_cap);
  if ((int64_t)C * BB > budget) C = (int32_t)(budget / BB);
  if ((int64_t)C * BB > rs.partials_cap) C = (int32_t)(rs.partials_cap / BB);
  if (C < 1) C = 1;
  k_tile_hist<<<dim3(ntiles, C), 512, (size_t)R * B * 4, st>>>(
      g.optr, g.oadj, g.ow, lab, vperm, vhist, W, N, B, R, C, C == 1 ? D : rs.partials);
  if (C == 1) {
  } else if (C <= 4)
    k_hist_reduce<1><<<grid_for(BB, T), T, 0, st>>>(rs.partials, C, BB, D);
  else if (C <= 32)
    k_hist_reduce<8><<<grid_for(BB * 8, T), T, 0, st>>>(rs.partials, C, BB, D);
  else
    k_hist_reduce<64><<<grid_for(BB * 64, T), T, 0, st>>>(rs.partials, C, BB, D);
  dense_to_csr(B, bm, rs, st);
}

static void small_path(const GraphDev &g, LabView lab, int32_t B, BMBuf &bm,
                       RebuildScratch &rs, hipStream_t st) {
  const int64_t BB = (int64_t)B * B;
  const int T = 256;
  int32_t *D = bm.dense;
  // edge chunks; about 2M partial counters in all (128 planes at B = 128)
  int32_t C = (int32_t)std::min<int64_t>(1024, std::max<int64_t>(128, ((int64_t)2 << 20) / BB));
  if ((int64_t)C * BB > rs.partials_cap) C = (int32_t)(rs.partials_cap / BB);
  if ((int64_t)C > g.E) C = (int32_t)std::max<int64_t>(g.E, 1);
  k_edge_hist<<<C, 512, (size_t)BB * 4, st>>>(g.esrc, g.oadj, g.ow, lab, g.E, B, C,
                                             C == 1 ? D : rs.partials);
  if (C == 1) {
  } else if (C <= 4)
    k_hist_reduce<1><<<grid_for(BB, T), T, 0, st>>>(rs.partials, C, BB, D);
  else if (C <= 32)
    k_hist_reduce<8><<<grid_for(BB * 8, T), T, 0, st>>>(rs.partials, C, BB, D);
  else
    k_hist_reduce<64><<<grid_for(BB * 64, T), T, 0, st>>>(rs.partials, C, BB, D);
  k_small_bm<<<1, 1024, 0, st>>>(D, B, bm.rowptr, bm.col, bm.val, bm.erow, bm.colptr, bm.row,
                                 bm.cval, bm.dout, bm.din, bm.ldout, bm.ldin);
}

// ---- incremental update of the dense matrix after a nodal batch
// One group of 16 lanes per vertex of the batch [lo, lo + BS); olab holds the batch's labels
// before the batch was applied.  Every edge with a moved endpoint leaves its old cell and
// enters its new one, counted once: the out-edges of a moved vertex (target labels old and
// new), and the in-edges of a moved vertex whose source did not move.
// (DT, when given, is D's transpose and receives the mirrored changes)
__global__ __launch_bounds__(256) void k_dense_moves(const int32_t *optr, const int32_t *oadj,
                                                     const int32_t *ow, const int32_t *iptr,
                                                     const int32_t *iadj, const int32_t *iw,
                                                     const int32_t *lab, const int32_t *olab,
                                                     int64_t lo, int64_t BS, int32_t B,
                                                     int32_t *D, int32_t *DT, int32_t *CS,
                                                     int32_t nch) {
  const int64_t i = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  const int q = threadIdx.x & 15;
  if (i >= BS) return;
  const int64_t x = lo + i;
  const int32_t xo = olab[i], xn = lab[x];
  if (xo == xn) return;
  for (int32_t j = optr[x] + q; j < optr[x + 1]; j += 16) {
    const int32_t y = oadj[j], w = ow[j];
    const int64_t k = (int64_t)y - lo;
    const int32_t yn = lab[y];
    const int32_t yo = (k >= 0 && k < BS) ? olab[k] : yn;
    atomicSub(&D[(int64_t)xo * B + yo], w);
    atomicAdd(&D[(int64_t)xn * B + yn], w);
    if (DT) {
      atomicSub(&DT[(int64_t)yo * B + xo], w);
      atomicAdd(&DT[(int64_t)yn * B + xn], w);
    }
    if (CS) {
      atomicSub(&CS[(int64_t)yo * nch + (xo >> 6)], w);
      atomicAdd(&CS[(int64_t)yn * nch + (xn >> 6)], w);
    }
  }
  for (int32_t j = iptr[x] + q; j < iptr[x + 1]; j += 16) {
    const int32_t y = iadj[j], w = iw[j];  // edge y -> x
    const int64_t k = (int64_t)y - lo;
    const int32_t yl = lab[y];
    if (k >= 0 && k < BS && olab[k] != yl) continue;  // y moved: one of y's out-edges
    atomicSub(&D[(int64_t)yl * B + xo], w);
    atomicAdd(&D[(int64_t)yl * B + xn], w);
    if (DT) {
      atomicSub(&DT[(int64_t)xo * B + yl], w);
      atomicAdd(&DT[(int64_t)xn * B + yl], w);
    }
    if (CS) {
      atomicSub(&CS[(int64_t)xo * nch + (yl >> 6)], w);
      atomicAdd(&CS[(int64_t)xn * nch + (yl >> 6)], w);
    }
  }
}

// Column chunk sums of the block matrix from its CSC: CS[x * nch + c] = the weight of
// column x in rows 64 c .. 64 c + 63.  A wave per column sums its entries into LDS and
// writes all nch words of the column (no clearing pass).
__global__ __launch_bounds__(256) void k_colsum_build(const int32_t *colptr, const int32_t *row,
                                                      const int32_t *cval, int32_t B, int32_t nch,
                                                      int32_t *CS) {
  extern __shared__ int32_t scs[];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int32_t *cs = scs + w * nch;
  const int32_t x = (int32_t)(blockIdx.x * 4 + w);
  for (int32_t c = lane; c < nch; c += 64) cs[c] = 0;
  wave_lds_sync();
  if (x < B)
    for (int32_t k = colptr[x] + lane; k < colptr[x + 1]; k += 64)
      atomicAdd(&cs[row[k] >> 6], cval[k]);
  wave_lds_sync();
  if (x < B)
    for (int32_t c = lane; c < nch; c += 64) CS[(int64_t)x * nch + c] = cs[c];
}

// B <= SMALL_B: the same edge moves summed in an LDS copy of the B x B matrix per workgroup,
// then added to D with one global atomic per changed cell: a few blocks would otherwise
// funnel every move into the same cells.  Each wave takes 64 vertices at a time (a lane
// each), finds the moved ones by ballot, and walks every moved vertex's edges with all 64
// lanes in turn (a vertex of thousands of neighbours is not left to a small team).
__global__ __launch_bounds__(512) void k_dense_moves_lds(const int32_t *optr, const int32_t *oadj,
                                                         const int32_t *ow, const int32_t *iptr,
                                                         const int32_t *iadj, const int32_t *iw,
                                                         const int32_t *lab, const int32_t *olab,
                                                         int64_t lo, int64_t BS, int32_t B,
                                                         int32_t *D) {
  extern __shared__ int32_t dl[];
  const int32_t BB = B * B;
  for (int32_t k = threadIdx.x; k < BB; k += 512) dl[k] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * 8;
  for (int64_t b0 = ((int64_t)blockIdx.x * 8 + (threadIdx.x >> 6)) * 64; b0 < BS; b0 += nw * 64) {
    const int64_t iv = b0 + lane;
    int32_t xo = 0, xn = 0;
    if (iv < BS) {
      xo = olab[iv];
      xn = lab[lo + iv];
    }
    uint64_t moved = __ballot(iv < BS && xo != xn);
    while (moved) {
      const int src = __ffsll((long long)moved) - 1;
      moved &= moved - 1;
      const int64_t x = lo + b0 + src;
      const int32_t vo = __shfl(xo, src, 64), vn = __shfl(xn, src, 64);
      const int32_t o0 = optr[x], o1 = optr[x + 1], i0 = iptr[x], i1 = iptr[x + 1];
      for (int32_t j = o0 + lane; j < o1; j += 64) {
        const int32_t y = oadj[j], w = ow[j];
        const int64_t k = (int64_t)y - lo;
        const int32_t yn = lab[y];
        const int32_t yo = (k >= 0 && k < BS) ? olab[k] : yn;
        atomicSub(&dl[vo * B + yo], w);
        atomicAdd(&dl[vn * B + yn], w);
      }
      for (int32_t j = i0 + lane; j < i1; j += 64) {
        const int32_t y = iadj[j], w = iw[j];  // edge y -> x
        const int64_t k = (int64_t)y - lo;
        const int32_t yl = lab[y];
        if (k >= 0 && k < BS && olab[k] != yl) continue;  // y moved: one of y's out-edges
        atomicSub(&dl[yl * B + vo], w);
        atomicAdd(&dl[yl * B + vn], w);
      }
    }
  }
  __syncthreads();
  for (int32_t k = threadIdx.x; k < BB; k += 512) {
    const int32_t d = dl[k];
    if (d != 0) atomicAdd(&D[k], d);
  }
}

// SMALL_B < B: workgroup (row tile, vertex chunk) sums the edge moves of its chunk's moved
// vertices that land in its MT_WORDS / B rows in an LDS copy of those rows, then adds them to
// D with one global atomic per changed cell.  Every chunk is read by every tile (the moved
// vertices are a small part of a batch), in exchange for no contention on hot cells.
constexpr int MT_WORDS = 16384;
__global__ __launch_bounds__(512) void k_dense_moves_tile(const int32_t *optr, const int32_t *oadj,
                                                          const int32_t *ow, const int32_t *iptr,
                                                          const int32_t *iadj, const int32_t *iw,
                                                          const int32_t *lab, const int32_t *olab,
                                                          int64_t lo, int64_t BS, int32_t B,
                                                          int32_t nchunk, int32_t *D) {
  __shared__ int32_t tl[MT_WORDS];
  __shared__ int32_t mv[512];
  __shared__ int32_t nmv;
  const int32_t RT = MT_WORDS / B;
  const int32_t tile = (int32_t)blockIdx.x / nchunk, chunk = (int32_t)blockIdx.x % nchunk;
  const int32_t r0 = tile * RT, r1 = min(B, r0 + RT);
  const int32_t cells = (r1 - r0) * B;
  for (int32_t k = threadIdx.x; k < cells; k += 512) tl[k] = 0;
  const int64_t c0 = BS * chunk / nchunk, c1 = BS * (chunk + 1) / nchunk;
  const int q = threadIdx.x & 15;
  // rounds of 512 vertices: the moved ones listed in LDS, then a group of 16 lanes each
  for (int64_t b0 = c0; b0 < c1; b0 += 512) {
    if (threadIdx.x == 0) nmv = 0;
    __syncthreads();
    const int64_t iv = b0 + threadIdx.x;
    if (iv < c1 && olab[iv] != lab[lo + iv]) mv[atomicAdd(&nmv, 1)] = (int32_t)(iv - b0);
    __syncthreads();
    const int32_t nm = nmv;
    for (int32_t m = threadIdx.x >> 4; m < nm; m += 32) {
      const int64_t i = b0 + mv[m];
      const int64_t x = lo + i;
      const int32_t xo = olab[i], xn = lab[x];
      const bool ho = xo >= r0 && xo < r1, hn = xn >= r0 && xn < r1;
      if (ho || hn) {
        for (int32_t j = optr[x] + q; j < optr[x + 1]; j += 16) {
          const int32_t y = oadj[j], w = ow[j];
          const int64_t k = (int64_t)y - lo;
          const int32_t yn = lab[y];
          const int32_t yo = (k >= 0 && k < BS) ? olab[k] : yn;
          if (ho) atomicSub(&tl[(xo - r0) * B + yo], w);
          if (hn) atomicAdd(&tl[(xn - r0) * B + yn], w);
        }
      }
      for (int32_t j = iptr[x] + q; j < iptr[x + 1]; j += 16) {
        const int32_t y = iadj[j], w = iw[j];
        const int64_t k = (int64_t)y - lo;
        const int32_t yl = lab[y];
        if (yl < r0 || yl >= r1 || (k >= 0 && k < BS && olab[k] != yl)) continue;
        atomicSub(&tl[(yl - r0) * B + xo], w);
        atomicAdd(&tl[(yl - r0) * B + xn], w);
      }
    }
    __syncthreads();  // the list is read before the next round rewrites it
  }
  __syncthreads();
  for (int32_t k = threadIdx.x; k < cells; k += 512) {
    const int32_t d = tl[k];
    if (d != 0) atomicAdd(&D[(int64_t)r0 * B + k], d);
  }
}


bool dense_update_ok(const BMBuf &bm, const RebuildScratch &rs, int32_t B) {
  return bm.dense_valid && bm.dense_B == B && bm.B == B && B <= INC_MAX_B &&
         (B > TAIL_MAX_B || (int64_t)B * B <= rs.partials_cap);
}

bool rebuild_after_moves(const GraphDev &g, const int32_t *lab32, const int32_t *olab, int64_t lo,
                         int64_t BS, int32_t B, BMBuf &bm, RebuildScratch &rs, hipStream_t st,
                         bool defer_csr) {
  // the dense buffer must hold the block matrix of the labels before the batch; the moves
  // are summed in an LDS copy of the matrix (B <= SMALL_B) or of a tile of its rows
  // (B <= INC_TILE_B, where global atomics would contend on few cells), or added with global
  // atomics (spread over many cells at larger B)
  if (!dense_update_ok(bm, rs, B)) return false;
  // above TAIL_MAX_B no transpose is kept: the column walks read D's column chunk sums,
  // built here from the CSC (current: the level's full rebuild or a sweep's last batch
  // preceded) and kept by the moves; below, the transpose follows the moves from the first
  // deferred batch on
  const bool cs_form = B > TAIL_MAX_B;
  const int32_t nch = (B + 63) / 64;
  if (cs_form && bm.cs_B != B) {
    if (bm.csr_stale || (int64_t)B * nch > rs.colsum_cap) return false;
    k_colsum_build<<<(unsigned)((B + 3) / 4), 256, (size_t)16 * nch, st>>>(
        bm.colptr, bm.row, bm.cval, B, nch, rs.colsum);
    bm.cs_B = B;
  } else if (!cs_form && defer_csr && B > INC_TILE_B && bm.dt_B != B) {
    const unsigned tb = (unsigned)((B + 31) / 32);
    k_transpose<<<dim3(tb, tb), 256, 0, st>>>(bm.dense, B, rs.partials);
    bm.dt_B = B;
  }
  if (BS > 0 && B <= SMALL_B)
    k_dense_moves_lds<<<(unsigned)std::min<int64_t>(512, (BS + 255) / 256), 512,
                        (size_t)B * B * 4, st>>>(
        g.optr, g.oadj, g.ow, g.iptr, g.iadj, g.iw, lab32, olab, lo, BS, B, bm.dense);
  else if (BS > 0 && B <= INC_TILE_B) {
    const int32_t tiles = (B + MT_WORDS / B - 1) / (MT_WORDS / B);
    const int32_t nchunk = (int32_t)std::max<int64_t>(1, std::min<int64_t>(512 / tiles, BS / 64));
    k_dense_moves_tile<<<(unsigned)(tiles * nchunk), 512, 0, st>>>(
        g.optr, g.oadj, g.ow, g.iptr, g.iadj, g.iw, lab32, olab, lo, BS, B, nchunk, bm.dense);
  }
  else if (BS > 0)  // the transpose (or the column chunk sums) follow when current
    k_dense_moves<<<(unsigned)((BS + 15) / 16), 256, 0, st>>>(
        g.optr, g.oadj, g.ow, g.iptr, g.iadj, g.iw, lab32, olab, lo, BS, B, bm.dense,
        !cs_form && bm.dt_B == B ? rs.partials : nullptr,
        cs_form && bm.cs_B == B ? rs.colsum : nullptr, nch);
  if (B <= INC_TILE_B) bm.dt_B = 0;  // the LDS forms change D alone
  if (defer_csr) {
    bm.csr_stale = true;
    return true;
  }
  // (the 16-bit label copy was kept current by the batch's apply)
  if (B <= SMALL_B)
    k_small_bm<<<1, 1024, 0, st>>>(bm.dense, B, bm.rowptr, bm.col, bm.val, bm.erow, bm.colptr,
                                   bm.row, bm.cval, bm.dout, bm.din, bm.ldout, bm.ldin);
  else if (B <= TAIL_MAX_B)
    dense_to_csr(B, bm, rs, st);
  else  // the dense tail reads 4 B^2 words: the segmented rebuild's CSR/CSC instead (D kept)
    seg_path(g, LabView{lab32, B <= 65535 ? rs.lab16 : nullptr}, B, bm, rs, false, st);
  bm.csr_stale = false;
  return true;
}

// ---- sparse update of the block matrix after a nodal batch (B without the dense copy)
// The batch's edge moves form a signed delta matrix dM.  For a moved vertex x (block xo
// before the batch, xn after): every out-edge x -> y of weight w gives row xo key yo: -w,
// row xn key yn: +w, column yo key xo: -w and column yn key xn: +w (yo, yn: y's labels
// before and after); every in-edge y -> x from a vertex y that did not move gives row y' key
// xo: -w, row y' key xn: +w, column xo key y': -w and column xn key y': +w (y': y's label).
// An edge between two moved vertices is counted once, as an out-edge.  Segment x < B of dM
// is row x and segment B + x column x, as in the segmented rebuild, whose sorts reduce each
// segment to runs (key, summed weight; the sums wrap in 32 bits, so negative totals come out
// in two's complement).  Each segment's runs D are then merged with the current row or
// column O into the alternate buffers (an entry whose sum is zero dropped) by a team sized
// to the segment: lane groups with a bitonic merge for up to 64 entries, a wave with the
// keys in LDS up to MW_CAP, a wave over 64-rank chunks above.  The work follows the moved
// vertices' edges plus one pass over the current entries, instead of sorting every edge
// again.

// nmv: the moved count followed by the merge classes' counts (8 words)
__global__ void k_delta_zero(uint32_t *cnt, uint32_t *cur, uint32_t *su, uint32_t *ncl,
                             uint32_t *nmv, int64_t nseg) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= nseg;
       i += (int64_t)gridDim.x * blockDim.x) {
    cnt[i] = 0;
    if (i < nseg) { cur[i] = 0; su[i] = 0; }
  }
  if (blockIdx.x == 0 && threadIdx.x < SEG_NCL) ncl[threadIdx.x] = 0;
  if (blockIdx.x == 0 && threadIdx.x < 8) nmv[threadIdx.x] = 0;
}

// the batch's moved vertices (index in the batch), listed in any order; *nmv their count
__global__ __launch_bounds__(256) void k_moved_list(const int32_t *lab, const int32_t *olab,
                                                    int64_t lo, int64_t BS, uint32_t *list,
                                                    uint32_t *nmv) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < BS; i0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = i0 + threadIdx.x;
    const bool mv = i < BS && olab[i] != lab[lo + i];
    const uint64_t m = __ballot(mv);
    uint32_t base = 0;
    if (lane == 0 && m) base = atomicAdd(nmv, (uint32_t)__popcll(m));
    base = __shfl(base, 0, 64);
    if (mv) list[base + (uint32_t)__popcll(m & lt)] = (uint32_t)i;
  }
}

// A group of 16 lanes per listed vertex walks its out-edges, then its in-edges.
// FILL = false: every delta counted into cnt[segment]; FILL = true: written at off[segment]
// plus a slot reserved in cur (a group reserves its chunk's entries of the vertex's own rows
// and columns with one atomic each).
template <bool FILL>
__global__ __launch_bounds__(256) void k_delta_gen(const int32_t *optr, const int32_t *oadj,
                                                   const int32_t *ow, const int32_t *iptr,
                                                   const int32_t *iadj, const int32_t *iw,
                                                   const int32_t *lab, const int32_t *olab,
                                                   int64_t lo, int64_t BS, const uint32_t *list,
                                                   const uint32_t *nmv, int32_t B,
                                                   uint32_t *cnt, const uint32_t *off,
                                                   uint32_t *cur, uint32_t *ekey, uint32_t *ew) {
  constexpr int G = 16;
  const int q = threadIdx.x & (G - 1);
  const int gbase = (threadIdx.x & 63) & ~(G - 1);
  const uint32_t lt = (1u << q) - 1u;
  const uint32_t Bu = (uint32_t)B;
  const uint32_t n = *nmv;
  const int64_t ng = (int64_t)gridDim.x * (blockDim.x / G);
  for (int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; k < n; k += ng) {
    const uint32_t i = list[k];
    const int64_t x = lo + i;
    const uint32_t vo = (uint32_t)olab[i], vn = (uint32_t)lab[x];
    const int32_t o0 = optr[x], o1 = optr[x + 1], i0 = iptr[x], i1 = iptr[x + 1];
    for (int32_t c = o0; c < o1; c += G) {
      const int32_t j = c + q;
      const bool on = j < o1;
      const uint32_t nc = (uint32_t)min(G, o1 - c);
      uint32_t ko = 0, kn = 0, w = 0;
      if (on) {
        const int32_t y = oadj[j];
        w = ow ? (uint32_t)ow[j] : 1u;
        const int64_t ky = (int64_t)y - lo;
        kn = (uint32_t)lab[y];
        ko = (ky >= 0 && ky < BS) ? (uint32_t)olab[ky] : kn;
      }
      if (!FILL) {
        if (q == 0) {
          atomicAdd(&cnt[vo], nc);
          atomicAdd(&cnt[vn], nc);
        }
        if (on) {
          if (ko == kn) {
            atomicAdd(&cnt[Bu + ko], 2u);
          } else {
            atomicAdd(&cnt[Bu + ko], 1u);
            atomicAdd(&cnt[Bu + kn], 1u);
          }
        }
      } else {
        uint32_t ro = 0, rn = 0;
        if (q == 0) {
          ro = off[vo] + atomicAdd(&cur[vo], nc);
          rn = off[vn] + atomicAdd(&cur[vn], nc);
        }
        ro = __shfl(ro, 0, G) + (uint32_t)q;
        rn = __shfl(rn, 0, G) + (uint32_t)q;
        if (on) {
          ekey[ro] = ko; ew[ro] = 0u - w;  // row vo, key yo: -w
          ekey[rn] = kn; ew[rn] = w;       // row vn, key yn: +w
          if (ko == kn) {
            const uint32_t p = off[Bu + ko] + atomicAdd(&cur[Bu + ko], 2u);
            ekey[p] = vo; ew[p] = 0u - w;
            ekey[p + 1] = vn; ew[p + 1] = w;
          } else {
            const uint32_t p = off[Bu + ko] + atomicAdd(&cur[Bu + ko], 1u);
            ekey[p] = vo; ew[p] = 0u - w;
            const uint32_t r = off[Bu + kn] + atomicAdd(&cur[Bu + kn], 1u);
            ekey[r] = vn; ew[r] = w;
          }
        }
      }
    }
    for (int32_t c = i0; c < i1; c += G) {
      const int32_t j = c + q;
      bool on = j < i1;
      uint32_t yl = 0, w = 0;
      if (on) {
        const int32_t y = iadj[j];  // edge y -> x
        w = iw ? (uint32_t)iw[j] : 1u;
        const int64_t ky = (int64_t)y - lo;
        yl = (uint32_t)lab[y];
        if (ky >= 0 && ky < BS && (uint32_t)olab[ky] != yl) on = false;  // y moved: its out-edge
      }
      const uint32_t m = (uint32_t)((__ballot(on) >> gbase) & 0xFFFFull);
      const uint32_t nc = (uint32_t)__popc(m);
      if (!FILL) {
        if (q == 0 && nc) {
          atomicAdd(&cnt[Bu + vo], nc);
          atomicAdd(&cnt[Bu + vn], nc);
        }
        if (on) atomicAdd(&cnt[yl], 2u);
      } else {
        uint32_t ro = 0, rn = 0;
        if (q == 0 && nc) {
          ro = off[Bu + vo] + atomicAdd(&cur[Bu + vo], nc);
          rn = off[Bu + vn] + atomicAdd(&cur[Bu + vn], nc);
        }
        const uint32_t r = (uint32_t)__popc(m & lt);
        ro = __shfl(ro, 0, G) + r;
        rn = __shfl(rn, 0, G) + r;
        if (on) {
          ekey[ro] = yl; ew[ro] = 0u - w;  // column vo, key y': -w
          ekey[rn] = yl; ew[rn] = w;       // column vn, key y': +w
          const uint32_t p = off[yl] + atomicAdd(&cur[yl], 2u);
          ekey[p] = vo; ew[p] = 0u - w;    // row y', key vo: -w
          ekey[p + 1] = vn; ew[p + 1] = w; // row y', key vn: +w
        }
      }
    }
  }
}

// Class C with B > SEG_RC (segments of more than SEG_G entries): the runs of each (segment,
// column range), counted by k_seg_c_count, written at the segment's start in rk / rw (the
// segment's own entries are re-read for every range, so they cannot be overwritten).
template <bool PK>
__global__ __launch_bounds__(512) void k_seg_c_runs(const uint32_t *off, const uint32_t *ekey,
                                                    const uint32_t *ew, const uint32_t *lists,
                                                    int64_t lcap, const uint32_t *ncl,
                                                    int32_t B, int32_t nr,
                                                    const uint32_t *cntc, uint32_t *rk,
                                                    uint32_t *rw, int wb) {
  extern __shared__ uint32_t cnt[];
  __shared__ uint32_t red8[8];
  const uint32_t RC = B < SEG_RC ? (uint32_t)B : (uint32_t)SEG_RC;
  const int64_t items = (int64_t)ncl[CL_C] * nr;
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int64_t k = it / nr;
    const uint32_t q = (uint32_t)(it - k * nr);
    if (cntc[it] == 0) continue;  // uniform across the workgroup
    const uint32_t x = lists[CL_C * lcap + k];
    uint32_t pre = 0;
    for (uint32_t qq = threadIdx.x; qq < q; qq += 512) pre += cntc[k * nr + qq];
    pre = block_sum512(pre, red8);
    const uint32_t c0 = q * RC, cn = min(RC, (uint32_t)B - c0);
    seg_range_count<PK>(cnt, ekey, ew, off[x], off[x + 1], c0, cn, wb);
    seg_emit(cnt, cn, c0, off[x] + pre, red8, rk, rw, nullptr, nullptr, nullptr, 0, nullptr);
  }
}

// The current matrix (read) and the alternate buffers (written) of a sparse update
struct BMOld {
  const int32_t *rowptr, *col, *val, *colptr, *row, *cval;
};
struct BMNew {
  int32_t *rowptr, *col, *val, *erow, *colptr, *row, *cval, *gcol, *grow;
  int64_t cap;  // entries each of col / val / erow / row / cval hold
};
// One segment of a sparse update: its current entries O (keys ascending) and delta runs D
// (keys ascending, signed weights)
struct MSeg {
  const int32_t *ok, *ov;
  const uint32_t *dk, *dv;
  int32_t m, d, y;
  bool isrow;
};
GSAP_D MSeg mseg(const BMOld &o, const uint32_t *off, const uint32_t *su, const uint32_t *ekey,
                 const uint32_t *ew, const uint32_t *rk, const uint32_t *rw, int one_range,
                 int32_t B, uint32_t x) {
  MSeg s;
  s.isrow = x < (uint32_t)B;
  s.y = (int32_t)(s.isrow ? x : x - B);
  const int32_t o0 = s.isrow ? o.rowptr[s.y] : o.colptr[s.y];
  s.m = (s.isrow ? o.rowptr[s.y + 1] : o.colptr[s.y + 1]) - o0;
  s.ok = (s.isrow ? o.col : o.row) + o0;
  s.ov = (s.isrow ? o.val : o.cval) + o0;
  const uint32_t s0 = off[x], nd = off[x + 1] - s0;
  s.d = nd ? (int32_t)su[x] : 0;
  const bool apart = !one_range && nd > (uint32_t)SEG_G;  // runs of the column-range counts
  s.dk = (apart ? rk : ekey) + s0;
  s.dv = (apart ? rw : ew) + s0;
  return s;
}
// output q of segment s (row or column entry, with its coarse-index word)
GSAP_D void mput(const BMNew &n, const MSeg &s, uint32_t q, int32_t key, int32_t v) {
  if ((int64_t)q >= n.cap) return;  // (a consistent matrix has at most E entries)
  if (s.isrow) {
    n.col[q] = key;
    n.val[q] = v;
    n.erow[q] = s.y;
    if (n.gcol && (q & 15u) == 0) n.gcol[q >> 4] = key;
  } else {
    n.row[q] = key;
    n.cval[q] = v;
    if (n.grow && (q & 15u) == 0) n.grow[q >> 4] = key;
  }
}
// first output position of segment x (rows, then columns from 0)
GSAP_D uint32_t mout0(const uint32_t *uoff, int32_t B, uint32_t x) {
  return x < (uint32_t)B ? uoff[x] : uoff[x] - uoff[B];
}

// merge classes by n = |O| + |D|: <= 16, 32, 64 (lane groups), <= MW_CAP (a wave with the
// keys in LDS), longer (a wave over 64-rank chunks)
enum { MC_16 = 0, MC_32 = 1, MC_64 = 2, MC_W = 3, MC_H = 4, MC_N = 5 };
constexpr int MW_CAP = 1024;

// Segments into class lists (empty ones get length 0 at once)
__global__ __launch_bounds__(256) void k_merge_classify(BMOld o, const uint32_t *off,
                                                        const uint32_t *su, int32_t B,
                                                        uint32_t *lists, int64_t lcap,
                                                        uint32_t *ncls, uint32_t *len) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t nseg = 2 * (int64_t)B;
  for (int64_t x0 = (int64_t)blockIdx.x * blockDim.x; x0 < nseg; x0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = x0 + threadIdx.x;
    int c = -1;
    if (x < nseg) {
      const bool isrow = x < B;
      const int32_t y = (int32_t)(isrow ? x : x - B);
      const int32_t m = isrow ? o.rowptr[y + 1] - o.rowptr[y] : o.colptr[y + 1] - o.colptr[y];
      const int32_t d = off[x + 1] != off[x] ? (int32_t)su[x] : 0;
      const int32_t n = m + d;
      if (n == 0) len[x] = 0;
      else c = n <= 16 ? MC_16 : n <= 32 ? MC_32 : n <= 64 ? MC_64 : n <= MW_CAP ? MC_W : MC_H;
    }
#pragma unroll
    for (int k = 0; k < MC_N; ++k) {
      const uint64_t msk = __ballot(c == k);
      if (!msk) continue;
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(&ncls[k], (uint32_t)__popcll(msk));
      base = __shfl(base, 0, 64);
      if (c == k) lists[k * lcap + base + (uint32_t)__popcll(msk & lt)] = (uint32_t)x;
    }
  }
}

// n <= G: a group of G lanes holds O ascending, padding, D descending (a bitonic sequence of
// (key << 1 | from-D) words) and merges it; an O entry and the D run of its key end up
// adjacent and are summed, zero values dropped, the rest compacted by ballot.
template <int G, bool WRITE>
GSAP_D void merge_group(const MSeg &s, uint32_t x, int32_t B, uint32_t *len, int32_t *dout,
                        int32_t *din, const BMNew &nw) {
  const int q = threadIdx.x & (G - 1);
  const int gbase = (threadIdx.x & 63) & ~(G - 1);
  const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
  uint32_t ck = 0xFFFFFFFFu;
  int32_t v = 0;
  if (q < s.m) {
    ck = (uint32_t)s.ok[q] << 1;
    v = s.ov[q];
  } else if (q >= G - s.d) {
    const int j = G - 1 - q;
    ck = (s.dk[j] << 1) | 1u;
    v = (int32_t)s.dv[j];
  }
  const int32_t dpart = (ck & 1u) && ck != 0xFFFFFFFFu ? v : 0;
#pragma unroll
  for (int j = G / 2; j > 0; j >>= 1) {
    const uint32_t pk = __shfl_xor(ck, j, G);
    const int32_t pv = __shfl_xor(v, j, G);
    const bool lower = (q & j) == 0;
    const bool take = lower ? pk < ck : pk > ck;
    if (take) { ck = pk; v = pv; }
  }
  const uint32_t nk = __shfl_down(ck, 1, G);
  const int32_t nv = __shfl_down(v, 1, G);
  const uint32_t pk = __shfl_up(ck, 1, G);
  const bool real = ck != 0xFFFFFFFFu;
  const bool head = real && !(ck & 1u) && q + 1 < G && (nk >> 1) == (ck >> 1) && nk != 0xFFFFFFFFu;
  const bool tail = real && (ck & 1u) && q > 0 && (pk >> 1) == (ck >> 1) && !(pk & 1u);
  const int32_t outv = head ? v + nv : v;
  const bool valid = real && !tail && outv != 0;
  const uint64_t mk = (__ballot(valid) >> gbase) & gmask;
  if (WRITE) {
    if (valid) {
      const uint32_t p = (uint32_t)__popcll(mk & ((1ull << q) - 1ull));
      mput(nw, s, mout0(len, B, x) + p, (int32_t)(ck >> 1), outv);
    }
  } else {
    int32_t ds = dpart;
#pragma unroll
    for (int o_ = G / 2; o_ >= 1; o_ >>= 1) ds += __shfl_xor(ds, o_, G);
    if (q == 0) {
      len[x] = (uint32_t)__popcll(mk);
      if (ds != 0) {
        if (s.isrow) dout[s.y] += ds;
        else din[s.y] += ds;
      }
    }
  }
}

// 64 < n <= MW_CAP: one wave, O and D keys staged in LDS (sk: O then D); each element's rank
// in the merged sequence comes from a binary search in the other list (ties: O first, so an
// O entry and its D run take consecutive ranks); sc[rank] = 1 for an output, scanned in place.
template <bool WRITE>
GSAP_D void merge_wave(const MSeg &s, uint32_t x, int32_t B, uint32_t *len, int32_t *dout,
                       int32_t *din, const BMNew &nw, int32_t *sk, uint32_t *sc) {
  const int lane = threadIdx.x & 63;
  const int32_t n = s.m + s.d;
  for (int32_t e = lane; e < n; e += 64) {
    sk[e] = e < s.m ? s.ok[e] : (int32_t)s.dk[e - s.m];
    sc[e] = 0;
  }
  wave_lds_sync();
  int32_t ds = 0;
  for (int32_t e = lane; e < n; e += 64) {
    const bool fromO = e < s.m;
    const int32_t key = sk[e];
    int32_t lo_ = 0, hi_ = fromO ? s.d : s.m;
    const int32_t *other = fromO ? sk + s.m : sk;
    while (lo_ < hi_) {
      const int32_t mid = (lo_ + hi_) >> 1;
      if (fromO ? other[mid] < key : other[mid] <= key) lo_ = mid + 1;
      else hi_ = mid;
    }
    bool valid;
    int32_t rank;
    if (fromO) {
      const bool pair = lo_ < s.d && other[lo_] == key;
      const int32_t v = s.ov[e] + (pair ? (int32_t)s.dv[lo_] : 0);
      valid = v != 0;
      rank = e + lo_;
    } else {
      const int32_t j = e - s.m;
      const bool pair = lo_ > 0 && other[lo_ - 1] == key;
      const int32_t w = (int32_t)s.dv[j];
      ds += w;
      valid = !pair && w != 0;
      rank = j + lo_;
    }
    if (valid) sc[rank] = 1;
  }
  wave_lds_sync();
  // exclusive scan of sc over the ranks
  uint32_t run = 0;
  for (int32_t b0 = 0; b0 < n; b0 += 64) {
    const int32_t r = b0 + lane;
    uint32_t f = r < n ? sc[r] : 0u, incl = f;
#pragma unroll
    for (int o_ = 1; o_ < 64; o_ <<= 1) {
      const uint32_t yv = __shfl_up(incl, o_, 64);
      if (lane >= o_) incl += yv;
    }
    if (r < n) sc[r] = run + incl - f;
    run += __shfl(incl, 63, 64);
  }
  wave_lds_sync();
  if (WRITE) {
    const uint32_t base = mout0(len, B, x);
    for (int32_t e = lane; e < n; e += 64) {
      const bool fromO = e < s.m;
      const int32_t key = sk[e];
      int32_t lo_ = 0, hi_ = fromO ? s.d : s.m;
      const int32_t *other = fromO ? sk + s.m : sk;
      while (lo_ < hi_) {
        const int32_t mid = (lo_ + hi_) >> 1;
        if (fromO ? other[mid] < key : other[mid] <= key) lo_ = mid + 1;
        else hi_ = mid;
      }
      if (fromO) {
        const bool pair = lo_ < s.d && other[lo_] == key;
        const int32_t v = s.ov[e] + (pair ? (int32_t)s.dv[lo_] : 0);
        if (v != 0) mput(nw, s, base + sc[e + lo_], key, v);
      } else {
        const int32_t j = e - s.m;
        const bool pair = lo_ > 0 && other[lo_ - 1] == key;
        const int32_t w = (int32_t)s.dv[j];
        if (!pair && w != 0) mput(nw, s, base + sc[j + lo_], key, w);
      }
    }
  } else {
#pragma unroll
    for (int o_ = 32; o_ >= 1; o_ >>= 1) ds += __shfl_xor(ds, o_, 64);
    if (lane == 0) {
      len[x] = run;
      if (ds != 0) {
        if (s.isrow) dout[s.y] += ds;
        else din[s.y] += ds;
      }
    }
  }
  wave_lds_sync();  // the LDS is reused by the wave's next segment
}

// Classes 16 / 32 / 64 / wave: wave tasks (four, two or one segments of a lane-group class,
// or one wave-class segment).  WRITE = false: len[x] = the segment's new length, and its runs'
// sum added to the block degree; WRITE = true (len scanned): the outputs, and the new row and
// column pointers.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_merge_small(BMOld o, const uint32_t *off,
                                                     const uint32_t *su, const uint32_t *ekey,
                                                     const uint32_t *ew, const uint32_t *rk,
                                                     const uint32_t *rw, int one_range,
                                                     int32_t B, const uint32_t *lists,
                                                     int64_t lcap, const uint32_t *ncls,
                                                     uint32_t *len, int32_t *dout, int32_t *din,
                                                     BMNew nw) {
  __shared__ int32_t sk[4][MW_CAP];
  __shared__ uint32_t sc[4][MW_CAP];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (WRITE) {
    const int64_t nseg = 2 * (int64_t)B;
    const uint32_t nnz = len[B];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= nseg;
         i += (int64_t)gridDim.x * blockDim.x) {
      if (i < B) nw.rowptr[i] = (int32_t)len[i];
      else if (i < nseg) nw.colptr[i - B] = (int32_t)(len[i] - nnz);
      else { nw.rowptr[B] = (int32_t)nnz; nw.colptr[B] = (int32_t)(len[nseg] - nnz); }
    }
  }
  const int64_t n16 = ncls[MC_16], n32 = ncls[MC_32], n64 = ncls[MC_64], nW = ncls[MC_W];
  const int64_t t16 = (n16 + 3) / 4, t32 = (n32 + 1) / 2;
  const int64_t tasks = t16 + t32 + n64 + nW;
  uint32_t *lenw = len;
  for (int64_t t = (int64_t)blockIdx.x * 4 + w; t < tasks; t += (int64_t)gridDim.x * 4) {
    if (t < t16) {
      const int64_t k = t * 4 + (lane >> 4);
      if (k < n16) {
        const uint32_t x = lists[MC_16 * lcap + k];
        const MSeg s = mseg(o, off, su, ekey, ew, rk, rw, one_range, B, x);
        merge_group<16, WRITE>(s, x, B, lenw, dout, din, nw);
      }
    } else if (t < t16 + t32) {
      const int64_t k = (t - t16) * 2 + (lane >> 5);
      if (k < n32) {
        const uint32_t x = lists[MC_32 * lcap + k];
        const MSeg s = mseg(o, off, su, ekey, ew, rk, rw, one_range, B, x);
        merge_group<32, WRITE>(s, x, B, lenw, dout, din, nw);
      }
    } else if (t < t16 + t32 + n64) {
      const uint32_t x = lists[MC_64 * lcap + (t - t16 - t32)];
      const MSeg s = mseg(o, off, su, ekey, ew, rk, rw, one_range, B, x);
      merge_group<64, WRITE>(s, x, B, lenw, dout, din, nw);
    } else {
      const uint32_t x = lists[MC_W * lcap + (t - t16 - t32 - n64)];
      const MSeg s = mseg(o, off, su, ekey, ew, rk, rw, one_range, B, x);
      merge_wave<WRITE>(s, x, B, lenw, dout, din, nw, sk[w], sc[w]);
    }
  }
}

// Class H (n > MW_CAP): one wave per listed segment merges O and D in key order, 64 outputs
// of the merged sequence at a time.  A chunk's split (a entries of O, b of D) is the merge
// path found by one ballot: lane t tests a = amin + 1 + t (O's a-th entry not above D's
// (L - a + 1)-th; ties put O first), and an equal (O, D) pair is kept in one chunk.  Each
// lane finds its rank in the chunk by a binary search over the other part (shuffles); an O
// entry adds its equal D run's weight, a D run without an equal O entry is new; entries of
// value 0 are dropped.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_merge_long(BMOld o, const uint32_t *off,
                                                    const uint32_t *su, const uint32_t *ekey,
                                                    const uint32_t *ew, const uint32_t *rk,
                                                    const uint32_t *rw, int one_range,
                                                    int32_t B, const uint32_t *lists,
                                                    int64_t lcap, const uint32_t *ncls,
                                                    uint32_t *len, int32_t *dout, int32_t *din,
                                                    BMNew nw) {
  const int lane = threadIdx.x & 63;
  const int64_t nl = ncls[MC_H];
  for (int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < nl; t += (int64_t)gridDim.x * 4) {
    const uint32_t x = lists[MC_H * lcap + t];
    const MSeg s = mseg(o, off, su, ekey, ew, rk, rw, one_range, B, x);
    const int32_t m = s.m, d = s.d;
    const uint32_t out0 = WRITE ? mout0(len, B, x) : 0u;
    int32_t io = 0, jd = 0;
    uint32_t pos = 0;
    int32_t dsum = 0;
    while (io < m || jd < d) {
      const int32_t remO = m - io, remD = d - jd;
      int32_t L = min(64, remO + remD);
      const int32_t amin = max(0, L - remD), amax = min(L, remO);
      // (a = amin always qualifies; lanes test amin + 1 .. amin + 64, which covers amax)
      const int32_t at = amin + 1 + lane;
      bool P = false;
      if (at <= amax) {
        if (L - at >= remD) P = true;
        else P = s.ok[io + at - 1] <= (int32_t)s.dk[jd + L - at];
      }
      int32_t a = amin + __popcll(__ballot(P));
      int32_t b = L - a;
      if (a > 0 && b < remD && s.ok[io + a - 1] == (int32_t)s.dk[jd + b]) {
        if (L < 64) { ++b; ++L; }
        else { --a; --L; }
      }
      const bool isO = lane < a, isD = lane >= a && lane < L;
      int32_t key = 0, v = 0;
      if (isO) {
        key = s.ok[io + lane];
        v = s.ov[io + lane];
      } else if (isD) {
        key = (int32_t)s.dk[jd + lane - a];
        v = (int32_t)s.dv[jd + lane - a];
      }
      // O entry: D runs of the chunk below its key; D run: O entries not above its key
      int32_t lo_ = 0, hi_ = isO ? b : (isD ? a : 0);
      const int base = isO ? a : 0;
#pragma unroll
      for (int st_ = 0; st_ < 7; ++st_) {
        const int32_t mid = (lo_ + hi_) >> 1;
        const int32_t kk = __shfl(key, (base + mid) & 63, 64);
        if (lo_ < hi_) {
          if (isO ? (kk < key) : (kk <= key)) lo_ = mid + 1;
          else hi_ = mid;
        }
      }
      const int32_t cnt = lo_;
      const int psrc = isO ? a + cnt : cnt - 1;
      const int32_t pk = __shfl(key, psrc & 63, 64), pv = __shfl(v, psrc & 63, 64);
      const bool matched = isO ? (cnt < b && pk == key) : (isD && cnt > 0 && pk == key);
      int32_t outv = v;
      bool valid;
      if (isO) {
        outv = v + (matched ? pv : 0);
        valid = outv != 0;
      } else {
        valid = isD && !matched && v != 0;
      }
      const int32_t rank = isO ? lane + cnt : lane - a + cnt;
      uint64_t mk = valid ? (1ull << rank) : 0ull;
#pragma unroll
      for (int q = 1; q < 64; q <<= 1) mk |= __shfl_xor(mk, q, 64);
      if (WRITE && valid)
        mput(nw, s, out0 + pos + (uint32_t)__popcll(mk & ((1ull << rank) - 1ull)), key, outv);
      if (!WRITE && isD) dsum += v;
      pos += (uint32_t)__popcll(mk);
      io += a;
      jd += b;
    }
    if (!WRITE) {
#pragma unroll
      for (int q = 32; q >= 1; q >>= 1) dsum += __shfl_xor(dsum, q, 64);
      if (lane == 0) {
        len[x] = pos;
        if (dsum != 0) {
          if (s.isrow) dout[s.y] += dsum;
          else din[s.y] += dsum;
        }
      }
    }
  }
}

bool rebuild_after_moves_sparse(const GraphDev &g, const int32_t *lab32, const int32_t *olab,
                                int64_t lo, int64_t BS, int32_t B, int64_t bound, BMBuf &bm,
                                RebuildScratch &rs, hipStream_t st) {
  if (!(rs.sparse_on && !bm.csr_stale && !bm.dense_valid && bm.dense_B == 0 && bm.B == B &&
        B > SMALL_B &&
        bm.a_rowptr && rs.rk && rs.mwork && bound <= rs.kcap && bound <= rs.rcap &&
        BS <= rs.mwork_list))
    return false;
  const int64_t nseg = 2 * (int64_t)B;
  const int T = 256;
  uint32_t *ekey = (uint32_t *)rs.k0, *ew = (uint32_t *)rs.k1;
  const bool one = B <= SEG_RC;
  const int32_t nr = (int32_t)((B + SEG_RC - 1) / SEG_RC);
  const uint32_t RC = one ? (uint32_t)B : (uint32_t)SEG_RC;
  const int32_t *ow = g.wmax == 1 ? nullptr : g.ow, *iw = g.wmax == 1 ? nullptr : g.iw;
  // the work area: the moved list, its count and the merge classes' counts
  uint32_t *list = rs.mwork, *nmv = list + rs.mwork_list, *mcls = nmv + 1;
  k_delta_zero<<<grid_for(nseg + 1, T, SEG_GRID), T, 0, st>>>(rs.soff, rs.vrel, rs.su, rs.ncl,
                                                             nmv, nseg);
  if (BS > 0) {
    k_moved_list<<<grid_for(BS, T, 1024), T, 0, st>>>(lab32, olab, lo, BS, list, nmv);
    const unsigned gg = (unsigned)std::min<int64_t>(4096, (BS + 15) / 16);
    k_delta_gen<false><<<gg, 256, 0, st>>>(g.optr, g.oadj, ow, g.iptr, g.iadj, iw, lab32, olab,
                                           lo, BS, list, nmv, B, rs.soff, nullptr, nullptr,
                                           nullptr, nullptr);
    scan_exclusive(rs.soff, rs.soff, nseg, rs.soff + nseg, rs.tmp, st);
    k_delta_gen<true><<<gg, 256, 0, st>>>(g.optr, g.oadj, ow, g.iptr, g.iadj, iw, lab32, olab,
                                          lo, BS, list, nmv, B, nullptr, rs.soff, rs.vrel, ekey,
                                          ew);
  } else {
    scan_exclusive(rs.soff, rs.soff, nseg, rs.soff + nseg, rs.tmp, st);
  }
  const uint32_t reg_max = (uint32_t)rs.seg_reg_max, count_min = (uint32_t)rs.seg_count_min;
  k_seg_classify<<<grid_for(nseg, T, SEG_GRID), T, 0, st>>>(rs.soff, nseg, one ? 1 : 0, reg_max,
                                                            count_min, rs.lists, rs.lcap, rs.ncl,
                                                            nullptr);
  if (one) {
    k_seg_c1<false><<<SEG_GRID, 512, (size_t)RC * 4, st>>>(rs.soff, ekey, ew, rs.lists, rs.lcap,
                                                           rs.ncl, B, rs.su, 0);
    if (count_min > (uint32_t)SEG_W) {
      k_seg_sort_g<16, CL_G4K, false><<<SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists,
                                                                rs.lcap, rs.ncl, rs.su, 0);
      k_seg_sort_g<8, CL_G2K, false><<<SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists,
                                                               rs.lcap, rs.ncl, rs.su, 0);
      k_seg_sort_g<4, CL_G1K, false><<<2 * SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists,
                                                                   rs.lcap, rs.ncl, rs.su, 0);
    }
  } else {
    k_seg_sort_g<16, CL_G4K, false><<<SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists,
                                                              rs.lcap, rs.ncl, rs.su, 0);
    k_seg_sort_g<8, CL_G2K, false><<<SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists,
                                                             rs.lcap, rs.ncl, rs.su, 0);
    k_seg_sort_g<4, CL_G1K, false><<<2 * SEG_GRID, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists,
                                                                 rs.lcap, rs.ncl, rs.su, 0);
    k_seg_c_count<false><<<SEG_GRID / 2, 512, (size_t)RC * 4, st>>>(
        rs.soff, ekey, ew, rs.lists, rs.lcap, rs.ncl, B, nr, rs.cntc, rs.su, 0);
    k_seg_c_runs<false><<<SEG_GRID / 2, 512, (size_t)RC * 4, st>>>(
        rs.soff, ekey, ew, rs.lists, rs.lcap, rs.ncl, B, nr, rs.cntc, rs.rk, rs.rw, 0);
  }
  k_seg_sort_w<false><<<SEG_GRID * 3 / 2, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists, rs.lcap,
                                                      rs.ncl, rs.su, 0);
  k_seg_sort_reg<false><<<SEG_GRID * 2, 256, 0, st>>>(rs.soff, ekey, ew, rs.lists, rs.lcap,
                                                      rs.ncl, rs.su, 0);
  // merge each segment's current entries with its delta runs: classes, new lengths (and
  // block degrees), offsets, outputs and pointers
  const BMOld o{bm.rowptr, bm.col, bm.val, bm.colptr, bm.row, bm.cval};
  const BMNew n{bm.a_rowptr, bm.a_col, bm.a_val, bm.a_erow, bm.a_colptr,
                bm.a_row, bm.a_cval, bm.a_gcol, bm.a_grow, g.E};
  k_merge_classify<<<grid_for(nseg, T, SEG_GRID), T, 0, st>>>(o, rs.soff, rs.su, B, rs.lists,
                                                              rs.lcap, mcls, rs.suoff);
  const int oc = one ? 1 : 0;
  k_merge_small<false><<<SEG_GRID * 2, 256, 0, st>>>(o, rs.soff, rs.su, ekey, ew, rs.rk, rs.rw,
                                                     oc, B, rs.lists, rs.lcap, mcls, rs.suoff,
                                                     bm.dout, bm.din, n);
  k_merge_long<false><<<SEG_GRID / 4, 256, 0, st>>>(o, rs.soff, rs.su, ekey, ew, rs.rk, rs.rw,
                                                    oc, B, rs.lists, rs.lcap, mcls, rs.suoff,
                                                    bm.dout, bm.din, n);
  scan_exclusive(rs.suoff, rs.suoff, nseg, rs.suoff + nseg, rs.tmp, st);
  k_merge_small<true><<<SEG_GRID * 2, 256, 0, st>>>(o, rs.soff, rs.su, ekey, ew, rs.rk, rs.rw,
                                                    oc, B, rs.lists, rs.lcap, mcls, rs.suoff,
                                                    bm.dout, bm.din, n);
  k_merge_long<true><<<SEG_GRID / 4, 256, 0, st>>>(o, rs.soff, rs.su, ekey, ew, rs.rk, rs.rw,
                                                   oc, B, rs.lists, rs.lcap, mcls, rs.suoff,
                                                   bm.dout, bm.din, n);
  k_logs<<<grid_for(B, 256), 256, 0, st>>>(bm.dout, bm.din, B, bm.ldout, bm.ldin);
  bm.swap_sets();
  return true;
}

void rebuild_blockmodel(const GraphDev &g, const int32_t *lab32, int32_t B, BMBuf &bm,
                        RebuildScratch &rs, int32_t dense_max_B, int32_t tile_max_B,
                        hipStream_t st) {
  const bool dense = B <= dense_max_B && (int64_t)B * B <= bm.dense_cap;
  const bool small = dense && B <= SMALL_B && (int64_t)B * B * 128 <= rs.partials_cap;
  // the row-tile path: at most TILE_MAX_B blocks (a tile of 16384 / B rows, at least two,
  // fills 64 KB of LDS); its vertex histogram (B x N / 2048 words) lives in the class lists
  // and is scanned with the shared scan temporary
  const int64_t hw = (int64_t)B * ((g.N + VT - 1) / VT);
  const bool tile = !small && dense && B <= tile_max_B && B <= TILE_MAX_B &&
                    (int64_t)B * B <= rs.partials_cap && hw <= SEG_NCL * rs.lcap &&
                    (int64_t)scan_temp_words(hw) <= rs.tmp_words;
  // restore the all-zero dense buffer, unless the small or tile path overwrites the same
  // B x B region in full
  if (bm.dense_B > 0 && !((small || tile) && bm.dense_B == B)) {
    k_dense_clear<<<grid_for(g.E, 256, SEG_GRID), 256, 0, st>>>(
        bm.erow, bm.col, bm.rowptr + bm.dense_B, bm.dense_B, bm.dense);
    bm.dense_B = 0;
  }
  bm.B = B;
  // the 16-bit label copy read here and by the next nodal batch (same labels)
  LabView lab{lab32, nullptr};
  if (B <= 65535 && rs.lab16) {
    k_lab16<<<grid_for(g.N, 256), 256, 0, st>>>(lab32, g.N, rs.lab16);
    lab.l16 = rs.lab16;
  }
  // the full paths rewrite D (and the small and tile paths the partial-plane buffer that
  // holds its transpose); dense_to_csr makes the transpose current again where it runs
  bm.dt_B = 0;
  if (small) small_path(g, lab, B, bm, rs, st);
  else if (tile) tile_path(g, lab, B, bm, rs, st);
  else seg_path(g, lab, B, bm, rs, dense, st);
  bm.dense_valid = dense;
  bm.dense_B = dense ? B : 0;
  bm.csr_stale = false;
  bm.cs_B = 0;  // D rewritten: the column chunk sums are rebuilt from the new CSC
  // lookups without the dense copy search long rows through the coarse index (the
  // segmented path, the only one without it, writes the index with the entries)
  bm.index_valid = !dense && bm.gcol && bm.grow;
}

// D = sum of the C partial planes of BB counters, then the CSR, CSC, degrees and their logs
// of the B <= 128 block matrix D (the small path's last two steps, on caller buffers)
void launch_small_from_planes(const int32_t *partials, int32_t C, int32_t B, int32_t *D,
                              int32_t *rowptr, int32_t *col, int32_t *val, int32_t *erow,
                              int32_t *colptr, int32_t *row, int32_t *cval, int32_t *dout,
                              int32_t *din, double *ldout, double *ldin, hipStream_t st) {
  const int64_t BB = (int64_t)B * B;
  const int T = 256;
  if (C <= 4)
    k_hist_reduce<1><<<grid_for(BB, T), T, 0, st>>>(partials, C, BB, D);
  else if (C <= 32)
    k_hist_reduce<8><<<grid_for(BB * 8, T), T, 0, st>>>(partials, C, BB, D);
  else
    k_hist_reduce<64><<<grid_for(BB * 64, T), T, 0, st>>>(partials, C, BB, D);
  k_small_bm<<<1, 1024, 0, st>>>(D, B, rowptr, col, val, erow, colptr, row, cval, dout, din,
                                 ldout, ldin);
}

}  // namespace gsap
