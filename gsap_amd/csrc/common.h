// Device/host primitives shared by every kernel of the SBP hot path.
//
//  * Philox4x32-10 counter-based RNG (Salmon et al., SC'11) replacing the clock-seeded
//    curand streams of GSAP/kernels.hpp:645-676, 859-874.
//  * Integer sampling helpers (mulhi, the random-vs-neighbour coin of kernels.hpp:720,
//    847, 1024 as an exact integer comparison).
//  * The canonical FP64 log2 and the MDL term w*(log2 w - (log2 a + log2 b)) with the
//    operand grouping of kernels.hpp:173-177.
//  * The canonical 64-way summation: 64 strided partials (position k goes to partial
//    k mod 64, ascending k), reduced by a binary tree with offsets 32..1.  A group of G
//    lanes (G | 64) evaluates it with 64/G partials per lane, so every group width
//    gives the same bits.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "log2_table.h"

#define GSAP_HD __host__ __device__ __forceinline__
#define GSAP_D __device__ __forceinline__

namespace gsap {

// ------------------------------------------------------------------ Philox
constexpr uint32_t PH_M0 = 0xD2511F53u, PH_M1 = 0xCD9E8D57u;
constexpr uint32_t PH_W0 = 0x9E3779B9u, PH_W1 = 0xBB67AE85u;
// counter word 3: purpose of the draw
constexpr uint32_t TAG_MERGE = 0x4D455247u;  // block-merge proposal (r, i)
constexpr uint32_t TAG_NODE = 0x4E4F4445u;   // nodal proposal of one vertex
constexpr uint32_t TAG_BLOCK = 0x424C4B53u;  // nodal per-block sample and coin
constexpr uint32_t TAG_GEN = 0x47454E52u;    // graph generator

struct Philox {
  uint32_t x0, x1, x2, x3;
};

GSAP_HD Philox philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                      uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    uint64_t p0 = (uint64_t)PH_M0 * c0;
    uint64_t p1 = (uint64_t)PH_M1 * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += PH_W0; k1 += PH_W1;
  }
  return Philox{c0, c1, c2, c3};
}

// uniform integer in [0, n) from a 32-bit draw (n < 2^32)
GSAP_HD uint32_t mulhi(uint32_t x, uint64_t n) { return (uint32_t)(((uint64_t)x * n) >> 32); }

// true with probability B/(deg+B): kernels.hpp:720 + 847 / 1024 as integers
GSAP_HD bool coin_random(uint32_t x, int64_t deg, int64_t B) {
  return (uint64_t)x * (uint64_t)(deg + B) < ((uint64_t)B << 32);
}

// ---------------------------------------------------------- canonical log2
// log2(x) = e + log2(c_j) + P(r),  x = 2^e m,  c_j = 1 + j/128 (top 7 mantissa bits),
// r = fma(m, 1/c_j, -1),  P(r) = sum_{k=1..9} (-1)^(k+1) r^k / (k ln 2) by Horner with fma.
static __constant__ double d_log2_inv[128] = {GSAP_LOG2_INV_VALUES};
static __constant__ double d_log2_lg[128] = {GSAP_LOG2_LG_VALUES};
static __constant__ double d_log2_poly[9] = {GSAP_LOG2_POLY_VALUES};
static const double h_log2_inv[128] = {GSAP_LOG2_INV_VALUES};
static const double h_log2_lg[128] = {GSAP_LOG2_LG_VALUES};
static const double h_log2_poly[9] = {GSAP_LOG2_POLY_VALUES};

GSAP_HD double clog2_core(double x, const double *inv, const double *lg, const double *poly) {
  if (x == 0.0) return -__builtin_huge_val();
  if (!(x > 0.0)) return __builtin_nan("");
  if (x == __builtin_huge_val()) return x;
  uint64_t bits = __builtin_bit_cast(uint64_t, x);
  int e = (int)((bits >> 52) & 0x7ff);
  if (e == 0) {
    double y = x * 18014398509481984.0;  // 2^54
    bits = __builtin_bit_cast(uint64_t, y);
    e = (int)((bits >> 52) & 0x7ff) - 54;
  }
  e -= 1023;
  int j = (int)((bits >> 45) & 127);
  double m = __builtin_bit_cast(double, (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  double r = __builtin_fma(m, inv[j], -1.0);
  double p = poly[8];
  p = __builtin_fma(p, r, poly[7]);
  p = __builtin_fma(p, r, poly[6]);
  p = __builtin_fma(p, r, poly[5]);
  p = __builtin_fma(p, r, poly[4]);
  p = __builtin_fma(p, r, poly[3]);
  p = __builtin_fma(p, r, poly[2]);
  p = __builtin_fma(p, r, poly[1]);
  p = __builtin_fma(p, r, poly[0]);
  p = p * r;
  return ((double)e + lg[j]) + p;
}

GSAP_D double clog2(double x) { return clog2_core(x, d_log2_inv, d_log2_lg, d_log2_poly); }
inline double clog2_host(double x) { return clog2_core(x, h_log2_inv, h_log2_lg, h_log2_poly); }

// w * (log2 w - (la + lb)), the term of kernels.hpp:173-177
GSAP_D double term(int64_t w, double la, double lb) {
  double dw = (double)w;
  return dw * (clog2(dw) - (la + lb));
}

// ----------------------------------------------------- FP32 log2 and terms
// Mixed-precision form: the same table split in float (entries rounded to nearest),
// log2(1+r) = r/ln2 (1 - r/2 + r^2/3 - r^3/4), |r| < 2^-7 (truncation < 2^-36).
// For x >= 1 (integer counts) the error against the exact log2 is at most
// (2 |log2 x| + 6) 2^-24, including the rounding of an integer x to float.
static __constant__ float d_log2_inv32[128] = {GSAP_LOG2_INV_VALUES};
static __constant__ float d_log2_lg32[128] = {GSAP_LOG2_LG_VALUES};
static const float h_log2_inv32[128] = {GSAP_LOG2_INV_VALUES};
static const float h_log2_lg32[128] = {GSAP_LOG2_LG_VALUES};

GSAP_HD float flog2_core(float x, const float *inv, const float *lg) {
  const uint32_t bits = __builtin_bit_cast(uint32_t, x);
  const int e = (int)((bits >> 23) & 0xff) - 127;  // x >= 1: normal
  const int j = (int)((bits >> 16) & 127);
  const float m = __builtin_bit_cast(float, (bits & 0x007FFFFFu) | 0x3F800000u);
  const float r = __builtin_fmaf(m, inv[j], -1.0f);
  float p = __builtin_fmaf(r, -0.25f, 0.33333334f);
  p = __builtin_fmaf(p, r, -0.5f);
  p = __builtin_fmaf(p, r, 1.0f);
  p = (p * r) * 1.44269504f;  // 1/ln 2
  return ((float)e + lg[j]) + p;
}
GSAP_D float flog2(float x) { return flog2_core(x, d_log2_inv32, d_log2_lg32); }
inline float flog2_host(float x) { return flog2_core(x, h_log2_inv32, h_log2_lg32); }

// FP32 term w * (log2 w - (la + lb)) and its error weight: with u = 2^-24 the rounding
// error of the returned value is at most 6 u w (|log2 w| + |la| + |lb| + 1); eb
// accumulates w (|log2 w| + |la| + |lb| + 1).
GSAP_D float term32(int64_t w, float la, float lb, float &eb) {
  const float fw = (float)w;
  const float L = flog2(fw);
  eb += fw * (((__builtin_fabsf(L) + __builtin_fabsf(la)) + __builtin_fabsf(lb)) + 1.0f);
  return fw * (L - (la + lb));
}

// Kahan-compensated float accumulator; combine() merges two of them through TwoSum.
struct Kahan {
  float s, c;
  GSAP_HD void init() { s = 0.0f; c = 0.0f; }
  GSAP_HD void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  GSAP_HD void combine(float s2, float c2) {
    const float t = s + s2;
    const float bp = t - s;
    const float err = (s - (t - bp)) + (s2 - bp);  // s + s2 = t + err exactly
    s = t;
    c = (c + c2) - err;
  }
  GSAP_HD float value() const { return s - c; }
};

// ------------------------------------------------------- canonical group sum
template <int G>
struct CAcc {
  static constexpr int S = 64 / G;
  double p[S];
  GSAP_D void init() {
#pragma unroll
    for (int j = 0; j < S; ++j) p[j] = 0.0;
  }
  // valid in the group's lane 0
  GSAP_D double reduce() {
#pragma unroll
    for (int off = 32; off >= G; off >>= 1) {
      const int d = off / G;
#pragma unroll
      for (int j = 0; j < S; ++j)
        if (j < d) p[j] += p[j + d];
    }
    double x = p[0];
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) x += __shfl_down(x, off, G);
    return x;
  }
};

// lanes of a group, lane index inside the group
template <int G>
GSAP_D int glane() {
  return (int)(threadIdx.x & (G - 1));
}

// ------------------------------------------------------------- hash tables
// Open addressing, linear probing; key -1 = empty.  Works on LDS or global
// memory through generic pointers; reads use atomic loads so that global
// tables never see stale L1 lines.
struct HTab {
  int32_t *key;
  int32_t *w;
  int32_t *first;
  uint32_t bits;
};

GSAP_D uint32_t ht_slot(int32_t t, uint32_t bits) {
  return bits == 0 ? 0u : (((uint32_t)t * 2654435761u) >> (32 - bits));
}
GSAP_D int32_t ld_relaxed(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GSAP_D void ht_clear(const HTab &h, int lane, int nl) {
  uint32_t cap = 1u << h.bits;
  for (uint32_t i = lane; i < cap; i += nl) {
    __hip_atomic_store(&h.key[i], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&h.w[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&h.first[i], 0x7fffffff, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
GSAP_D void ht_insert(const HTab &h, int32_t t, int32_t w, int32_t pos) {
  uint32_t mask = (1u << h.bits) - 1u;
  uint32_t s = ht_slot(t, h.bits);
  for (;;) {
    int32_t old = atomicCAS(&h.key[s], -1, t);
    if (old == -1 || old == t) {
      atomicAdd(&h.w[s], w);
      atomicMin(&h.first[s], pos);
      return;
    }
    s = (s + 1) & mask;
  }
}
// slot of key t, or -1
GSAP_D int32_t ht_find(const HTab &h, int32_t t) {
  uint32_t mask = (1u << h.bits) - 1u;
  uint32_t s = ht_slot(t, h.bits);
  for (;;) {
    int32_t k = ld_relaxed(&h.key[s]);
    if (k == t) return (int32_t)s;
    if (k == -1) return -1;
    s = (s + 1) & mask;
  }
}
GSAP_D int32_t ht_weight(const HTab &h, int32_t t) {
  int32_t s = ht_find(h, t);
  return s < 0 ? 0 : ld_relaxed(&h.w[s]);
}

// ------------------------------------------------------------------ labels
// Block labels as read by the gather-heavy kernels: a 16-bit copy when every label fits
// (half the cache footprint of the int32 array; the int32 array of a 2M-vertex graph no
// longer fits one XCD's L2), else the int32 array.
struct LabView {
  const int32_t *l32;
  const uint16_t *l16;
  GSAP_D int32_t operator[](int64_t i) const { return l16 ? (int32_t)l16[i] : l32[i]; }
};

// ------------------------------------------------------------ block matrix
// Device view of the block matrix M (CSR + CSC, degrees, cached logs).
struct BMView {
  int32_t B;
  const int32_t *rowptr, *col, *val;     // CSR, columns ascending
  const int32_t *colptr, *row, *cval;    // CSC, rows ascending
  const int32_t *dout, *din;
  const double *ldout, *ldin;
  const int32_t *dense;                  // B*B row-major, or nullptr
  // without the dense copy: every 16th entry of col / row (gcol[j] = col[16 j]), a coarse
  // index that turns a long row's binary search into a few lines (nullptr: plain search)
  const int32_t *gcol, *grow;
};

// Position of t in the sorted keys[lo, hi), or -1.  Long ranges first search the coarse
// index (idx[j] = keys[16 j]) for the 16-entry block that can hold t, then that block.
GSAP_D int32_t idx_find(const int32_t *keys, const int32_t *idx, int32_t lo, int32_t hi,
                        int32_t t) {
  int32_t a = lo, b = hi;
  if (idx && hi - lo > 32) {
    const int32_t jl = (lo + 15) >> 4, jh = ((hi - 1) >> 4) + 1;
    int32_t l = jl, h = jh;  // first indexed entry greater than t
    while (l < h) {
      const int32_t mid = (l + h) >> 1;
      if (idx[mid] <= t) l = mid + 1;
      else h = mid;
    }
    if (l == jl) {
      b = jl << 4;
    } else {
      a = (l - 1) << 4;
      b = a + 16 < hi ? a + 16 : hi;
    }
  }
  while (a < b) {
    const int32_t mid = (a + b) >> 1;
    if (keys[mid] < t) a = mid + 1;
    else b = mid;
  }
  return (a < hi && keys[a] == t) ? a : -1;
}

// Both positions of t: in keys[lo1, hi1) and keys[lo2, hi2) (two rows of one array),
// the two searches advancing in the same loop so their loads overlap.
GSAP_D void idx_find2(const int32_t *keys, const int32_t *idx, int32_t lo1, int32_t hi1,
                      int32_t lo2, int32_t hi2, int32_t t, int32_t &k1, int32_t &k2) {
  int32_t a1 = lo1, b1 = hi1, a2 = lo2, b2 = hi2;
  if (idx) {
    const bool u1 = hi1 - lo1 > 32, u2 = hi2 - lo2 > 32;
    const int32_t j1 = (lo1 + 15) >> 4, j2 = (lo2 + 15) >> 4;
    int32_t l1 = j1, h1 = u1 ? ((hi1 - 1) >> 4) + 1 : j1;
    int32_t l2 = j2, h2 = u2 ? ((hi2 - 1) >> 4) + 1 : j2;
    while (l1 < h1 || l2 < h2) {
      if (l1 < h1) {
        const int32_t mid = (l1 + h1) >> 1;
        if (idx[mid] <= t) l1 = mid + 1;
        else h1 = mid;
      }
      if (l2 < h2) {
        const int32_t mid = (l2 + h2) >> 1;
        if (idx[mid] <= t) l2 = mid + 1;
        else h2 = mid;
      }
    }
    if (u1) {
      if (l1 == j1) b1 = j1 << 4;
      else { a1 = (l1 - 1) << 4; b1 = a1 + 16 < hi1 ? a1 + 16 : hi1; }
    }
    if (u2) {
      if (l2 == j2) b2 = j2 << 4;
      else { a2 = (l2 - 1) << 4; b2 = a2 + 16 < hi2 ? a2 + 16 : hi2; }
    }
  }
  while (a1 < b1 || a2 < b2) {
    if (a1 < b1) {
      const int32_t mid = (a1 + b1) >> 1;
      if (keys[mid] < t) a1 = mid + 1;
      else b1 = mid;
    }
    if (a2 < b2) {
      const int32_t mid = (a2 + b2) >> 1;
      if (keys[mid] < t) a2 = mid + 1;
      else b2 = mid;
    }
  }
  k1 = (a1 < hi1 && keys[a1] == t) ? a1 : -1;
  k2 = (a2 < hi2 && keys[a2] == t) ? a2 : -1;
}

// M[r][t] (0 when absent)
GSAP_D int32_t bm_row_get(const BMView &m, int32_t r, int32_t t) {
  if (m.dense) return m.dense[(int64_t)r * m.B + t];
  const int32_t k = idx_find(m.col, m.gcol, m.rowptr[r], m.rowptr[r + 1], t);
  return k >= 0 ? m.val[k] : 0;
}
// M[t][s] via column s (0 when absent)
GSAP_D int32_t bm_col_get(const BMView &m, int32_t s, int32_t t) {
  if (m.dense) return m.dense[(int64_t)t * m.B + s];
  const int32_t k = idx_find(m.row, m.grow, m.colptr[s], m.colptr[s + 1], t);
  return k >= 0 ? m.cval[k] : 0;
}

// M[r][t] and M[s][t] together
GSAP_D void bm_row_get2(const BMView &m, int32_t r, int32_t s, int32_t t, int32_t &Mr,
                        int32_t &Ms) {
  if (m.dense) {
    Mr = m.dense[(int64_t)r * m.B + t];
    Ms = m.dense[(int64_t)s * m.B + t];
    return;
  }
  int32_t k1, k2;
  idx_find2(m.col, m.gcol, m.rowptr[r], m.rowptr[r + 1], m.rowptr[s], m.rowptr[s + 1], t, k1,
            k2);
  Mr = k1 >= 0 ? m.val[k1] : 0;
  Ms = k2 >= 0 ? m.val[k2] : 0;
}

// M[t][r] and M[t][s] together via columns r and s
GSAP_D void bm_col_get2(const BMView &m, int32_t r, int32_t s, int32_t t, int32_t &Mr,
                        int32_t &Ms) {
  if (m.dense) {
    Mr = m.dense[(int64_t)t * m.B + r];
    Ms = m.dense[(int64_t)t * m.B + s];
    return;
  }
  int32_t k1, k2;
  idx_find2(m.row, m.grow, m.colptr[r], m.colptr[r + 1], m.colptr[s], m.colptr[s + 1], t, k1,
            k2);
  Mr = k1 >= 0 ? m.cval[k1] : 0;
  Ms = k2 >= 0 ? m.cval[k2] : 0;
}

// Weighted pick over row x then column x: first entry whose running weight
// exceeds target, skipping id `skip` (its mass carries forward); `fallback` if
// none qualifies.  Sequential (kernels.hpp:767-816, 978-995 as integers).
GSAP_D int32_t bm_walk(const BMView &m, int32_t x, uint64_t target, int32_t skip,
                       int32_t fallback) {
  uint64_t cum = 0;
  for (int32_t k = m.rowptr[x]; k < m.rowptr[x + 1]; ++k) {
    cum += (uint64_t)m.val[k];
    if (cum > target && m.col[k] != skip) return m.col[k];
  }
  for (int32_t k = m.colptr[x]; k < m.colptr[x + 1]; ++k) {
    cum += (uint64_t)m.cval[k];
    if (cum > target && m.row[k] != skip) return m.row[k];
  }
  return fallback;
}

// Orders the LDS operations of one wave (LDS that is private to the wave).
GSAP_D void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Total order on doubles as unsigned keys (a radix sort on float bits orders so).
GSAP_HD uint64_t order_key(double d) {
  uint64_t u = __builtin_bit_cast(uint64_t, d);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

}  // namespace gsap
