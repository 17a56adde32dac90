// Block-matrix rebuild: labels -> M as CSR (columns ascending) + CSC (rows ascending),
// degrees and cached log2 of the degrees.  Replaces calculate_new_csr,
// GSAP/partition.hpp:707-849 (sort_by_key of vertices, node-level gathers, two bb_segsort
// calls, run flags, remove_if, scans and mgpu::segreduce).
//
// Two paths, chosen by B:
//   dense   (B <= dense_max_B): M is histogrammed into a B*B int32 matrix -- per-workgroup
//           LDS histograms + a column-parallel reduction when B*B fits 64 KB of LDS, global
//           atomics otherwise -- then compacted row-wise (wave ballots) and column-wise.
//           The dense matrix stays valid for O(1) M[r][t] lookups by the sweep kernels.
//   sparse  (larger B): key = b(u)*B + b(v) per edge, stable radix sort, run-length encode
//           (integer atomics, exact), then a second radix sort of (col, row) keys for the CSC.
#include <algorithm>

#include "engine.h"

namespace gsap {

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---- vertices grouped by block (counting sort; order inside a block is irrelevant)
constexpr int VT = 2048;  // vertices per workgroup

__global__ __launch_bounds__(256) void k_label_hist(const int32_t *lab, int64_t N, int32_t B,
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

__global__ __launch_bounds__(256) void k_label_scatter(const int32_t *lab, int64_t N, int32_t B,
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
                                                   const int32_t *ow, const int32_t *lab,
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

// B <= 128: the whole B x B histogram in LDS; workgroup c takes a contiguous chunk of the
// edges (id order, one edge per thread per step, so that many label gathers are in flight)
// and writes its plane of partial counts (or D itself when C == 1).
__global__ __launch_bounds__(512) void k_edge_hist(const int32_t *esrc, const int32_t *oadj,
                                                   const int32_t *ow, const int32_t *lab,
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

// one wave per row: nonzero count and row sum
__global__ __launch_bounds__(64) void k_dense_rows(const int32_t *D, int32_t B, uint32_t *rcnt,
                                                   int32_t *dout) {
  const int32_t r = blockIdx.x;
  const int lane = threadIdx.x;
  uint32_t c = 0;
  int32_t s = 0;
  for (int32_t t = lane; t < B; t += 64) {
    int32_t x = D[(int64_t)r * B + t];
    c += x != 0;
    s += x;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    c += __shfl_down(c, off, 64);
    s += __shfl_down(s, off, 64);
  }
  if (lane == 0) { rcnt[r] = c; dout[r] = s; }
}

__global__ __launch_bounds__(64) void k_dense_write_rows(const int32_t *D, int32_t B,
                                                         const int32_t *rowptr, int32_t *col,
                                                         int32_t *val, int32_t *erow) {
  const int32_t r = blockIdx.x;
  const int lane = threadIdx.x;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int32_t run = rowptr[r];
  for (int32_t base = 0; base < B; base += 64) {
    int32_t t = base + lane;
    int32_t x = t < B ? D[(int64_t)r * B + t] : 0;
    uint64_t m = __ballot(x != 0);
    if (x != 0) {
      int32_t pos = run + (int32_t)__popcll(m & lt);
      col[pos] = t; val[pos] = x;
      if (erow) erow[pos] = r;
    }
    run += (int32_t)__popcll(m);
  }
}

// one wave per segment: out[x] = sum of vals[ptr[x] .. ptr[x+1])
__global__ __launch_bounds__(256) void k_seg_sum(const int32_t *ptr, const int32_t *vals,
                                                 int32_t B, int32_t *out) {
  const int32_t x = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (x >= B) return;
  const int lane = threadIdx.x & 63;
  int32_t s = 0;
  for (int32_t k = ptr[x] + lane; k < ptr[x + 1]; k += 64) s += vals[k];
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) out[x] = s;
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

__global__ void k_logs(const int32_t *dout, const int32_t *din, int32_t B, double *ldout,
                       double *ldin) {
  const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= B) return;
  ldout[x] = clog2((double)dout[x]);
  ldin[x] = clog2((double)din[x]);
}

// Small block matrices (B <= 128): CSR, CSC, degrees and their logs from D in one
// workgroup of 1024 threads (16 waves).  Row r (column c) is one wave's ballot compaction
// of D's row (column) in ascending order, as k_dense_write_rows does.
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

// ---------------------------------------------------------------- sparse path
template <class K>
__global__ void k_make_keys(const int32_t *esrc, const int32_t *oadj, const int32_t *ow,
                            const int32_t *lab, int64_t E, int32_t B, K *keys, uint32_t *vals) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    keys[e] = (K)lab[esrc[e]] * (K)B + (K)lab[oadj[e]];
    vals[e] = (uint32_t)ow[e];
  }
}

template <class K>
__global__ void k_run_flags(const K *keys, int64_t E, uint32_t *flags) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x)
    flags[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1u : 0u;
}

template <class K>
__global__ void k_rle(const K *keys, const uint32_t *vals, const uint32_t *flags,
                      const uint32_t *uidx, int64_t E, int32_t B, int32_t *col, int32_t *val,
                      int32_t *erow) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x) {
    // uidx is the exclusive scan of the run-head flags: a run's index is
    // uidx + flag - 1 for every element of the run
    uint32_t i = uidx[e] + flags[e] - 1u;
    atomicAdd(&val[i], (int32_t)vals[e]);
    if (flags[e]) {
      K k = keys[e];
      col[i] = (int32_t)(k % (K)B);
      erow[i] = (int32_t)(k / (K)B);
    }
  }
}

// ptr[x] = first entry whose major index (maj[i]) >= x; ptr[B] already holds nnz
__global__ void k_ptr_from_major(const int32_t *maj, const int32_t *nnz_dev, int32_t B,
                                 int32_t *ptr) {
  const int32_t nnz = *nnz_dev;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (nnz == 0) {
    for (int64_t x = i; x < B; x += (int64_t)gridDim.x * blockDim.x) ptr[x] = 0;
    return;
  }
  for (int64_t k = i; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
    int32_t r = maj[k];
    int32_t prev = k == 0 ? -1 : maj[k - 1];
    for (int32_t x = prev + 1; x <= r; ++x) ptr[x] = (int32_t)k;
    if (k == nnz - 1)
      for (int32_t x = r + 1; x < B; ++x) ptr[x] = nnz;
  }
}

// CSC keys: the column of every CSR entry (entries past nnz get B and sort last); the
// value is the CSR position.  A stable sort by column of the row-major CSR order leaves
// rows ascending inside each column.
__global__ void k_make_ckeys(const int32_t *col, const int32_t *nnz_dev, int64_t n, int32_t B,
                             uint32_t *keys, uint32_t *vals) {
  const int32_t nnz = *nnz_dev;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    keys[k] = k < nnz ? (uint32_t)col[k] : (uint32_t)B;
    vals[k] = (uint32_t)k;
  }
}

__global__ void k_csc_gather(const uint32_t *keys, const uint32_t *vals, const int32_t *nnz_dev,
                             const int32_t *erow, const int32_t *val, int32_t *row,
                             int32_t *cval, int32_t *ccol) {
  const int32_t nnz = *nnz_dev;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t i = vals[k];
    row[k] = erow[i];
    cval[k] = val[i];
    ccol[k] = (int32_t)keys[k];
  }
}

__global__ void k_copy_i32(const int32_t *a, int32_t *b) { *b = *a; }

static unsigned grid_for(int64_t n, int threads, unsigned cap = 65535u * 4u) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > (int64_t)cap) g = cap;
  return (unsigned)g;
}

static int bits_for(uint64_t x) {  // number of bits needed to represent values < x
  int b = 0;
  while (b < 64 && (x - 1) >> b) ++b;
  return b;
}

template <class K>
static void sparse_path(const GraphDev &g, const int32_t *lab, int32_t B, BMBuf &bm,
                        RebuildScratch &rs, hipStream_t st) {
  const int64_t E = g.E;
  const int T = 256;
  K *k0 = (K *)rs.k0, *k1 = (K *)rs.k1;
  uint32_t *v0 = rs.v0, *v1 = rs.v1;
  const int bits = bits_for((uint64_t)B * (uint64_t)B);
  int32_t *nnz_dev = bm.rowptr + B;
  k_make_keys<K><<<grid_for(E, T), T, 0, st>>>(g.esrc, g.oadj, g.ow, lab, E, B, k0, v0);
  radix_sort_pairs<K>(k0, v0, k1, v1, E, bits, rs.tmp, st);
  k_run_flags<K><<<grid_for(E, T), T, 0, st>>>(k0, E, rs.flags);
  scan_exclusive(rs.flags, rs.uidx, E, (uint32_t *)nnz_dev, rs.tmp, st);
  k_fill_i32<<<grid_for(E, T), T, 0, st>>>(bm.val, E, 0);
  k_rle<K><<<grid_for(E, T), T, 0, st>>>(k0, v0, rs.flags, rs.uidx, E, B, bm.col, bm.val,
                                         bm.erow);
  k_ptr_from_major<<<grid_for(E, T), T, 0, st>>>(bm.erow, nnz_dev, B, bm.rowptr);
  k_seg_sum<<<(unsigned)((B + 3) / 4), 256, 0, st>>>(bm.rowptr, bm.val, B, bm.dout);
  // CSC: stable sort of the CSR entries by column alone
  uint32_t *c0 = (uint32_t *)rs.k0, *c1 = (uint32_t *)rs.k1;
  v0 = rs.v0; v1 = rs.v1;
  k_make_ckeys<<<grid_for(E, T), T, 0, st>>>(bm.col, nnz_dev, E, B, c0, v0);
  radix_sort_pairs<uint32_t>(c0, v0, c1, v1, E, bits_for((uint64_t)B + 1), rs.tmp, st);
  int32_t *ccol = (int32_t *)rs.flags;  // reuse: column of each CSC entry
  k_csc_gather<<<grid_for(E, T), T, 0, st>>>(c0, v0, nnz_dev, bm.erow, bm.val, bm.row, bm.cval,
                                            ccol);
  k_copy_i32<<<1, 1, 0, st>>>(nnz_dev, bm.colptr + B);
  k_ptr_from_major<<<grid_for(E, T), T, 0, st>>>(ccol, nnz_dev, B, bm.colptr);
  k_seg_sum<<<(unsigned)((B + 3) / 4), 256, 0, st>>>(bm.colptr, bm.cval, B, bm.din);
}

static void dense_path(const GraphDev &g, const int32_t *lab, int32_t B, BMBuf &bm,
                       RebuildScratch &rs, hipStream_t st) {
  const int64_t N = g.N;
  const int64_t BB = (int64_t)B * B;
  const int T = 256;
  int32_t *D = bm.dense;
  // row tiles of at most 16384 counters (64 KB of LDS), chunks so that ~2048 workgroups run
  const int32_t R = B >= 16384 ? 1 : 16384 / B;
  const int32_t ntiles = (B + R - 1) / R;
  const int32_t W = (int32_t)((N + VT - 1) / VT);
  const bool one_tile = ntiles == 1;
  if (!one_tile) {  // vertices grouped by block, so that a tile reads only its rows' vertices
    k_label_hist<<<W, 256, (size_t)B * 4, st>>>(lab, N, B, rs.vhist, W);
    scan_exclusive(rs.vhist, rs.vhist, (int64_t)B * W, nullptr, rs.tmp, st);
    k_label_scatter<<<W, 256, (size_t)B * 4, st>>>(lab, N, B, rs.vhist, W, rs.vperm);
  }
  int32_t C;
  if (one_tile) {
    // edge chunks; about 2M partial counters in all (128 planes at B = 128)
    C = (int32_t)std::min<int64_t>(1024, std::max<int64_t>(128, ((int64_t)2 << 20) / BB));
    if ((int64_t)C * BB > rs.partials_cap) C = (int32_t)(rs.partials_cap / BB);
    if ((int64_t)C > g.E) C = (int32_t)std::max<int64_t>(g.E, 1);
    k_edge_hist<<<C, 512, (size_t)BB * 4, st>>>(g.esrc, g.oadj, g.ow, lab, g.E, B, C,
                                               C == 1 ? D : rs.partials);
  } else {
    C = 2048 / ntiles;
    if ((int64_t)C * BB > ((int64_t)8 << 20)) C = (int32_t)(((int64_t)8 << 20) / BB);
    if ((int64_t)C * BB > rs.partials_cap) C = (int32_t)(rs.partials_cap / BB);
    if (C < 1) C = 1;
    // a single chunk writes its tiles straight into D
    k_tile_hist<<<dim3(ntiles, C), 512, (size_t)R * B * 4, st>>>(
        g.optr, g.oadj, g.ow, lab, rs.vperm, rs.vhist, W, N, B, R, C, C == 1 ? D : rs.partials);
  }
  if (C == 1) {
  } else if (C <= 4)
    k_hist_reduce<1><<<grid_for(BB, T), T, 0, st>>>(rs.partials, C, BB, D);
  else if (C <= 32)
    k_hist_reduce<8><<<grid_for(BB * 8, T), T, 0, st>>>(rs.partials, C, BB, D);
  else
    k_hist_reduce<64><<<grid_for(BB * 64, T), T, 0, st>>>(rs.partials, C, BB, D);
  if (B <= SMALL_B) {
    k_small_bm<<<1, 1024, 0, st>>>(D, B, bm.rowptr, bm.col, bm.val, bm.erow, bm.colptr, bm.row,
                                   bm.cval, bm.dout, bm.din, bm.ldout, bm.ldin);
    return;
  }
  k_dense_rows<<<B, 64, 0, st>>>(D, B, rs.cnt, bm.dout);
  scan_exclusive(rs.cnt, (uint32_t *)bm.rowptr, B, (uint32_t *)(bm.rowptr + B), rs.tmp, st);
  k_dense_write_rows<<<B, 64, 0, st>>>(D, B, bm.rowptr, bm.col, bm.val, bm.erow);
  // columns: the same row kernels on the transpose (partials are free again here)
  int32_t *DT = rs.partials;
  const unsigned tb = (unsigned)((B + 31) / 32);
  k_transpose<<<dim3(tb, tb), 256, 0, st>>>(D, B, DT);
  k_dense_rows<<<B, 64, 0, st>>>(DT, B, rs.cnt, bm.din);
  scan_exclusive(rs.cnt, (uint32_t *)bm.colptr, B, (uint32_t *)(bm.colptr + B), rs.tmp, st);
  k_dense_write_rows<<<B, 64, 0, st>>>(DT, B, bm.colptr, bm.row, bm.cval, nullptr);
}

void rebuild_blockmodel(const GraphDev &g, const int32_t *lab, int32_t B, BMBuf &bm,
                        RebuildScratch &rs, int32_t dense_max_B, hipStream_t st) {
  bm.B = B;
  const bool dense = B <= dense_max_B && (int64_t)B * B <= bm.dense_cap &&
                     (int64_t)B * B <= rs.partials_cap && B <= rs.vhist_B;
  if (dense) {
    dense_path(g, lab, B, bm, rs, st);
    if (B <= SMALL_B) {  // k_small_bm wrote the logs
      bm.dense_valid = true;
      return;
    }
  } else if ((uint64_t)B * (uint64_t)B <= 0xFFFFFFFFull) {
    sparse_path<uint32_t>(g, lab, B, bm, rs, st);
  } else {
    sparse_path<uint64_t>(g, lab, B, bm, rs, st);
  }
  bm.dense_valid = dense;
  k_logs<<<grid_for(B, 256), 256, 0, st>>>(bm.dout, bm.din, B, bm.ldout, bm.ldin);
}

}  // namespace gsap
