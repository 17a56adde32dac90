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
#include "engine.h"

namespace gsap {

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_hist_lds(const int32_t *esrc, const int32_t *oadj, const int32_t *ow,
                           const int32_t *lab, int64_t E, int32_t B, int32_t *partials) {
  extern __shared__ int32_t h[];
  const int32_t BB = B * B;
  for (int32_t i = threadIdx.x; i < BB; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const int64_t chunk = (E + gridDim.x - 1) / gridDim.x;
  const int64_t e0 = (int64_t)blockIdx.x * chunk;
  const int64_t e1 = e0 + chunk < E ? e0 + chunk : E;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x)
    atomicAdd(&h[lab[esrc[e]] * B + lab[oadj[e]]], ow[e]);
  __syncthreads();
  int32_t *out = partials + (int64_t)blockIdx.x * BB;
  for (int32_t i = threadIdx.x; i < BB; i += blockDim.x) out[i] = h[i];
}

__global__ void k_hist_reduce(const int32_t *partials, int32_t nparts, int32_t BB, int32_t *D) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= BB) return;
  int32_t s = 0;
  for (int32_t p = 0; p < nparts; ++p) s += partials[(int64_t)p * BB + i];
  D[i] = s;
}

__global__ void k_hist_global(const int32_t *esrc, const int32_t *oadj, const int32_t *ow,
                              const int32_t *lab, int64_t E, int32_t B, int32_t *D) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&D[(int64_t)lab[esrc[e]] * B + lab[oadj[e]]], ow[e]);
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

template <class K>
__global__ void k_make_tkeys(const int32_t *col, const int32_t *erow, const int32_t *val,
                             const int32_t *nnz_dev, int64_t n, int32_t B, K pad, K *keys,
                             uint32_t *vals) {
  const int32_t nnz = *nnz_dev;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    if (k < nnz) {
      keys[k] = (K)col[k] * (K)B + (K)erow[k];
      vals[k] = (uint32_t)val[k];
    } else {
      keys[k] = pad;
      vals[k] = 0;
    }
  }
}

template <class K>
__global__ void k_csc_unpack(const K *keys, const uint32_t *vals, const int32_t *nnz_dev,
                             int32_t B, int32_t *row, int32_t *cval, int32_t *ccol) {
  const int32_t nnz = *nnz_dev;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * blockDim.x) {
    K key = keys[k];
    row[k] = (int32_t)(key % (K)B);
    ccol[k] = (int32_t)(key / (K)B);
    cval[k] = (int32_t)vals[k];
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
  // CSC: sort (col, row) keys of the nnz entries; padding sorts behind them
  const K pad = (bits >= (int)(8 * sizeof(K))) ? (K)~(K)0 : (K)(((K)1 << bits) - 1);
  k0 = (K *)rs.k0; k1 = (K *)rs.k1; v0 = rs.v0; v1 = rs.v1;
  k_make_tkeys<K><<<grid_for(E, T), T, 0, st>>>(bm.col, bm.erow, bm.val, nnz_dev, E, B, pad, k0,
                                                v0);
  radix_sort_pairs<K>(k0, v0, k1, v1, E, bits, rs.tmp, st);
  int32_t *ccol = (int32_t *)rs.flags;  // reuse: column of each CSC entry
  k_csc_unpack<K><<<grid_for(E, T), T, 0, st>>>(k0, v0, nnz_dev, B, bm.row, bm.cval, ccol);
  k_copy_i32<<<1, 1, 0, st>>>(nnz_dev, bm.colptr + B);
  k_ptr_from_major<<<grid_for(E, T), T, 0, st>>>(ccol, nnz_dev, B, bm.colptr);
  k_seg_sum<<<(unsigned)((B + 3) / 4), 256, 0, st>>>(bm.colptr, bm.cval, B, bm.din);
}

static void dense_path(const GraphDev &g, const int32_t *lab, int32_t B, BMBuf &bm,
                       RebuildScratch &rs, hipStream_t st) {
  const int64_t E = g.E;
  const int32_t BB = B * B;
  const int T = 256;
  int32_t *D = bm.dense;
  const int nparts = 256;
  if ((int64_t)BB * 4 <= 65536 && (int64_t)nparts * BB <= rs.partials_cap) {
    k_hist_lds<<<nparts, 512, (size_t)BB * 4, st>>>(g.esrc, g.oadj, g.ow, lab, E, B,
                                                     rs.partials);
    k_hist_reduce<<<grid_for(BB, T), T, 0, st>>>(rs.partials, nparts, BB, D);
  } else {
    k_fill_i32<<<grid_for(BB, T), T, 0, st>>>(D, BB, 0);
    k_hist_global<<<grid_for(E, T, 8192), T, 0, st>>>(g.esrc, g.oadj, g.ow, lab, E, B, D);
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
                     (int64_t)B * B <= rs.partials_cap;
  if (dense) {
    dense_path(g, lab, B, bm, rs, st);
  } else if ((uint64_t)B * (uint64_t)B <= 0xFFFFFFFFull) {
    sparse_path<uint32_t>(g, lab, B, bm, rs, st);
  } else {
    sparse_path<uint64_t>(g, lab, B, bm, rs, st);
  }
  bm.dense_valid = dense;
  k_logs<<<grid_for(B, 256), 256, 0, st>>>(bm.dout, bm.din, B, bm.ldout, bm.ldin);
}

}  // namespace gsap
