// Segmented-free primitives for CDNA4: device-wide exclusive scan and a stable LSD
// radix sort of (key, value) pairs.  They replace the thrust scans / sort_by_key and
// the bb_segsort calls of GSAP/partition.hpp:726-825 and 442-445.
//
// Radix sort: 8-bit digits, tiles of 256 threads x 16 items.  Per pass:
//   1. per-tile digit histogram (LDS atomics) -> hist[digit * ntiles + tile]
//   2. exclusive scan of the digit-major histogram -> global scatter offsets
//   3. stable scatter: items are ranked in rounds of 256 (one item per thread); inside a
//      wave64 the rank comes from 8 ballots (peers with equal digit) + popcount, across the
//      four waves of the round from an LDS prefix, across rounds from running digit bases.
#include "prims.h"

namespace gsap {

constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;

GSAP_D uint32_t wave_incl_scan(uint32_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  return x;
}

// Block-wide exclusive scan of one value per thread (256 threads = 4 waves).
GSAP_D uint32_t block_excl_scan(uint32_t x, uint32_t *lds4, uint32_t *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl = wave_incl_scan(x);
  if (lane == 63) lds4[w] = incl;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t c = lds4[i];
    if (i < w) off += c;
    tot += c;
  }
  __syncthreads();
  if (total) *total = tot;
  return off + incl - x;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const uint32_t *in, int64_t n,
                                                        uint32_t *tile_sums) {
  __shared__ uint32_t lds4[4];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_I;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_I; ++i)
    if (base + i < n) s += in[base + i];
  uint32_t tot;
  block_excl_scan(s, lds4, &tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// single block: exclusive scan of ntiles values in place, total to *total
__global__ __launch_bounds__(SCAN_T) void k_scan_tiles(uint32_t *sums, int64_t ntiles,
                                                       uint32_t *total) {
  __shared__ uint32_t lds4[4];
  uint32_t carry = 0;
  for (int64_t c = 0; c < ntiles; c += SCAN_T) {
    int64_t i = c + threadIdx.x;
    uint32_t x = i < ntiles ? sums[i] : 0u;
    uint32_t tot;
    uint32_t ex = block_excl_scan(x, lds4, &tot);
    if (i < ntiles) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_apply(const uint32_t *in, uint32_t *out,
                                                       int64_t n, const uint32_t *tile_offs) {
  __shared__ uint32_t lds4[4];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_I;
  uint32_t v[SCAN_I];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_I; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0u;
    s += v[i];
  }
  uint32_t ex = block_excl_scan(s, lds4, nullptr) + tile_offs[blockIdx.x];
#pragma unroll
  for (int i = 0; i < SCAN_I; ++i) {
    if (base + i < n) out[base + i] = ex;
    ex += v[i];
  }
}

// single block, whole array: tiles in order with a running carry (small n)
__global__ __launch_bounds__(SCAN_T) void k_scan_single(const uint32_t *in, uint32_t *out,
                                                        int64_t n, uint32_t *total) {
  __shared__ uint32_t lds4[4];
  uint32_t carry = 0;
  for (int64_t t0 = 0; t0 < n; t0 += SCAN_TILE) {
    const int64_t base = t0 + (int64_t)threadIdx.x * SCAN_I;
    uint32_t v[SCAN_I];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_I; ++i) {
      v[i] = (base + i < n) ? in[base + i] : 0u;
      s += v[i];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan(s, lds4, &tot) + carry;
#pragma unroll
    for (int i = 0; i < SCAN_I; ++i) {
      if (base + i < n) out[base + i] = ex;
      ex += v[i];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

// single workgroup of 1024 threads, tiles of 8192 elements in order (n up to a few tiles)
constexpr int SCAN_T1K = 1024, SCAN_TILE1K = SCAN_T1K * SCAN_I;
__global__ __launch_bounds__(SCAN_T1K) void k_scan_single1k(const uint32_t *in, uint32_t *out,
                                                            int64_t n, uint32_t *total) {
  __shared__ uint32_t lds16[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t carry = 0;
  for (int64_t t0 = 0; t0 < n; t0 += SCAN_TILE1K) {
    const int64_t base = t0 + (int64_t)threadIdx.x * SCAN_I;
    uint32_t v[SCAN_I];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_I; ++i) {
      v[i] = (base + i < n) ? in[base + i] : 0u;
      s += v[i];
    }
    const uint32_t incl = wave_incl_scan(s);
    if (lane == 63) lds16[w] = incl;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const uint32_t c = lds16[i];
      if (i < w) pre += c;
      tot += c;
    }
    __syncthreads();
    uint32_t ex = carry + pre + incl - s;
#pragma unroll
    for (int i = 0; i < SCAN_I; ++i) {
      if (base + i < n) out[base + i] = ex;
      ex += v[i];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

size_t scan_temp_words(int64_t n) { return (size_t)((n + SCAN_TILE - 1) / SCAN_TILE) + 1; }

void scan_exclusive(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total, uint32_t *tmp,
                    hipStream_t st) {
  int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nt == 0) {
    if (total) hipMemsetAsync(total, 0, 4, st);
    return;
  }
  if (nt <= 2) {  // up to 4096 elements: one launch
    k_scan_single<<<1, SCAN_T, 0, st>>>(in, out, n, total);
    return;
  }
  if (n <= 2 * SCAN_TILE1K) {  // up to 16384 elements: one launch of 1024 threads
    k_scan_single1k<<<1, SCAN_T1K, 0, st>>>(in, out, n, total);
    return;
  }
  k_scan_reduce<<<(unsigned)nt, SCAN_T, 0, st>>>(in, n, tmp);
  k_scan_tiles<<<1, SCAN_T, 0, st>>>(tmp, nt, total);
  k_scan_apply<<<(unsigned)nt, SCAN_T, 0, st>>>(in, out, n, tmp);
}

// ------------------------------------------------------------------ radix sort
constexpr int RS_T = 256, RS_I = 16, RS_TILE = RS_T * RS_I;

template <class K>
__global__ __launch_bounds__(RS_T) void k_radix_hist(const K *keys, int64_t n, int shift,
                                                     uint32_t *hist, int64_t ntiles) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
  for (int it = 0; it < RS_I; ++it) {
    int64_t e = base + it * RS_T + threadIdx.x;
    if (e < n) atomicAdd(&cnt[(uint32_t)(keys[e] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter of one tile: items are ranked in rounds of 256 (one per thread; inside a
// wave64 from 8 ballots + popcount, across the four waves from an LDS prefix, across rounds
// from running digit bases), placed in LDS in tile-local digit order, then written out in
// that order so that each digit's run goes to consecutive addresses.  The tile's count of
// digit d is the step between its scanned offset and the next entry of the digit-major
// offset array.
template <class K>
__global__ __launch_bounds__(RS_T) void k_radix_scatter(const K *kin, const uint32_t *vin,
                                                        K *kout, uint32_t *vout, int64_t n,
                                                        int shift, const uint32_t *offs,
                                                        int64_t ntiles) {
  __shared__ K sk[RS_TILE];
  __shared__ uint32_t sv[RS_TILE];
  __shared__ uint32_t goff[256], loff[256], run[256];
  __shared__ uint32_t wc[4][256];
  __shared__ uint32_t lds4[4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  {
    const int64_t idx = (int64_t)tid * ntiles + blockIdx.x;
    const uint32_t g = offs[idx];
    const uint32_t nx = (idx + 1 < 256 * ntiles) ? offs[idx + 1] : (uint32_t)n;
    goff[tid] = g;
    const uint32_t ex = block_excl_scan(nx - g, lds4, nullptr);
    loff[tid] = ex;
    run[tid] = ex;
  }
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t tbase = (int64_t)blockIdx.x * RS_TILE;
  for (int it = 0; it < RS_I; ++it) {
    const int64_t e = tbase + it * RS_T + tid;
    const bool valid = e < n;
    K k = valid ? kin[e] : (K)0;
    uint32_t v = valid ? vin[e] : 0u;
    const uint32_t d = (uint32_t)(k >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
#pragma unroll
    for (int i = 0; i < 4; ++i) wc[i][tid] = 0;
    __syncthreads();
    if (valid && (peers & lt) == 0) wc[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    {
      uint32_t c0 = wc[0][tid], c1 = wc[1][tid], c2 = wc[2][tid], c3 = wc[3][tid];
      wc[0][tid] = 0;
      wc[1][tid] = c0;
      wc[2][tid] = c0 + c1;
      wc[3][tid] = c0 + c1 + c2;
      __syncthreads();
      if (valid) {
        const uint32_t pos = run[d] + wc[w][d] + rank;
        sk[pos] = k;
        sv[pos] = v;
      }
      __syncthreads();
      run[tid] += c0 + c1 + c2 + c3;
    }
    __syncthreads();
  }
  const int64_t left = n - tbase;
  const int32_t tn = left < RS_TILE ? (int32_t)left : RS_TILE;
  for (int32_t i = tid; i < tn; i += RS_T) {
    const K k = sk[i];
    const uint32_t d = (uint32_t)(k >> shift) & 255u;
    const uint32_t pos = goff[d] + ((uint32_t)i - loff[d]);
    kout[pos] = k;
    vout[pos] = sv[i];
  }
}

size_t radix_temp_words(int64_t n) {
  int64_t nt = (n + RS_TILE - 1) / RS_TILE;
  return (size_t)(256 * nt) + scan_temp_words(256 * nt) + 8;
}

template <class K>
int radix_sort_pairs(K *&keys, uint32_t *&vals, K *&keys_alt, uint32_t *&vals_alt, int64_t n,
                     int bits, uint32_t *tmp, hipStream_t st) {
  if (n <= 1 || bits <= 0) return 0;
  int64_t nt = (n + RS_TILE - 1) / RS_TILE;
  uint32_t *hist = tmp;
  uint32_t *stmp = tmp + 256 * nt;
  int passes = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    k_radix_hist<K><<<(unsigned)nt, RS_T, 0, st>>>(keys, n, shift, hist, nt);
    scan_exclusive(hist, hist, 256 * nt, nullptr, stmp, st);
    k_radix_scatter<K><<<(unsigned)nt, RS_T, 0, st>>>(keys, vals, keys_alt, vals_alt, n, shift,
                                                        hist, nt);
    K *tk = keys; keys = keys_alt; keys_alt = tk;
    uint32_t *tv = vals; vals = vals_alt; vals_alt = tv;
    ++passes;
  }
  return passes;
}

template int radix_sort_pairs<uint32_t>(uint32_t *&, uint32_t *&, uint32_t *&, uint32_t *&,
                                        int64_t, int, uint32_t *, hipStream_t);
template int radix_sort_pairs<uint64_t>(uint64_t *&, uint32_t *&, uint64_t *&, uint32_t *&,
                                        int64_t, int, uint32_t *, hipStream_t);

}  // namespace gsap
