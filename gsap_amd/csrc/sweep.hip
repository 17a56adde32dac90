// Kernels of one SBP level: per-block terms and samples, nodal moves, block merges,
// reductions.  Every FP64 sum follows the canonical 64-way order of common.h, so the
// results are bit-identical for any group width and match the CPU oracle.
#include "engine.h"

namespace gsap {

// --------------------------------------------------------------- helpers
GSAP_D double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_down(x, off, 64);
  return x;  // lane 0
}

// LDS copies of the canonical log2 tables: every lookup of a kernel that loads them at its
// start is an LDS read instead of a vector-cache read (same values, so the same bits)
__shared__ double s_log2_inv[128], s_log2_lg[128];
GSAP_D void log2_tables_to_lds() {
  for (int i = threadIdx.x; i < 128; i += blockDim.x) {
    s_log2_inv[i] = d_log2_inv[i];
    s_log2_lg[i] = d_log2_lg[i];
  }
  __syncthreads();
}
GSAP_D double clog2_s(double x) { return clog2_core(x, s_log2_inv, s_log2_lg, d_log2_poly); }
// term() of common.h with the LDS tables
GSAP_D double term_s(int64_t w, double la, double lb) {
  const double dw = (double)w;
  return dw * (clog2_s(dw) - (la + lb));
}

GSAP_D uint32_t wave_incl_scan_u32(uint32_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  return x;
}

// Weighted pick over v's out list then in list: the first entry whose running weight
// exceeds target (kernels.hpp:916-936 as integers).  Wave-wide; every lane returns it.
GSAP_D int32_t wave_pick(const int32_t *w1, const int32_t *id1, int32_t n1, const int32_t *w2,
                         const int32_t *id2, int32_t n2, uint32_t target) {
  constexpr int U = 4;  // chunks whose weights and ids are loaded together
  const int lane = threadIdx.x & 63;
  uint32_t base = 0;
  const int32_t n = n1 + n2;
  for (int32_t c = 0; c < n; c += U * 64) {
    uint32_t w[U];
    int32_t id[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t k = c + 64 * u + lane;
      w[u] = 0;
      id[u] = 0;
      if (k < n1) { w[u] = (uint32_t)w1[k]; id[u] = id1[k]; }
      else if (k < n) { w[u] = (uint32_t)w2[k - n1]; id[u] = id2[k - n1]; }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t incl = wave_incl_scan_u32(w[u]);
      const uint64_t mask = __ballot(base + incl > target);
      if (mask) return __shfl(id[u], __ffsll((long long)mask) - 1, 64);
      base += __shfl(incl, 63, 64);
    }
  }
  return id1 ? id1[0] : id2[0];  // unreachable when target < total weight
}

// The same pick by a whole 256-thread workgroup: thread t sums a contiguous run of
// ceil(n / 256) entries, a block scan of the run sums finds the run holding the target, and
// that run's thread finishes the search.  A constant number of dependent memory rounds
// instead of one per 64 entries.  sc: 5 words of LDS.  Every thread returns the id.
GSAP_D int32_t block_pick(const int32_t *w1, const int32_t *id1, int32_t n1, const int32_t *w2,
                          const int32_t *id2, int32_t n2, uint32_t target, uint32_t *sc) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int32_t n = n1 + n2;
  const int32_t K = (n + 255) / 256;
  const int32_t a0 = tid * K, a1 = a0 + K < n ? a0 + K : n;
  uint32_t sum = 0;
  for (int32_t k = a0; k < a1; ++k) sum += (uint32_t)(k < n1 ? w1[k] : w2[k - n1]);
  uint32_t incl = wave_incl_scan_u32(sum);
  if (lane == 63) sc[w] = incl;
  if (tid == 0) sc[4] = (uint32_t)(n1 > 0 ? id1[0] : id2[0]);  // unreachable fallback
  __syncthreads();
  uint32_t pre = 0;
  for (int i = 0; i < w; ++i) pre += sc[i];
  const uint32_t ex = pre + incl - sum;
  __syncthreads();
  if (ex <= target && target - ex < sum) {  // the one run where the running weight crosses
    uint32_t run = ex;
    for (int32_t k = a0; k < a1; ++k) {
      run += (uint32_t)(k < n1 ? w1[k] : w2[k - n1]);
      if (run > target) {
        sc[4] = (uint32_t)(k < n1 ? id1[k] : id2[k - n1]);
        break;
      }
    }
  }
  __syncthreads();
  return (int32_t)sc[4];
}

template <int G, class F>
GSAP_D void seg_for(int32_t n, F &&f) {
  const int q = threadIdx.x & (G - 1);
  for (int32_t base = 0; base < n; base += 64) {
#pragma unroll
    for (int j = 0; j < 64 / G; ++j) {
      const int32_t k = base + q + j * G;
      if (k < n) f(k, j);
    }
  }
}

// seg_for with the loads batched: P entries per lane at a time (in the lane's entry order),
// first a(k) for all P (the entry's own loads; k clamped into [0, n) so every load is in
// bounds), then b(a(k)) for all P (loads that depend on the entry), then c(a, b, j) for the
// entries below n in order.  Per slot j the entries arrive in the order seg_for gives them,
// so canonical sums come out the same; the batching only lets the loads overlap.
// an entry's looked-up count and log2 degree
struct GL {
  int32_t g;
  double l;
};

template <int G, int P, class A, class Bf, class C>
GSAP_D void seg_for3(int32_t n, A &&a, Bf &&b, C &&c) {
  constexpr int S = 64 / G;                   // entries per lane in a 64-entry chunk
  constexpr int CH = P >= S ? P / S : 1;      // chunks per step
  constexpr int NB = P >= S ? 1 : S / P;      // batches per step
  const int q = threadIdx.x & (G - 1);
  using TA = decltype(a(0));
  using TB = decltype(b(a(0)));
  for (int32_t base = 0; base < n; base += 64 * CH) {
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
      TA x[P];
      TB y[P];
#pragma unroll
      for (int e = 0; e < P; ++e) {
        const int el = bb * P + e;
        const int32_t k = base + 64 * (el / S) + q + (el % S) * G;
        x[e] = a(k < n ? k : n - 1);
      }
#pragma unroll
      for (int e = 0; e < P; ++e) y[e] = b(x[e]);
#pragma unroll
      for (int e = 0; e < P; ++e) {
        const int el = bb * P + e;
        const int32_t k = base + 64 * (el / S) + q + (el % S) * G;
        if (k < n) c(x[e], y[e], el % S);
      }
    }
  }
}

GSAP_D int64_t wave_sum_i64(int64_t x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_down(x, off, 64);
  return x;  // lane 0
}

static unsigned grid_for(int64_t n, int threads) {
  int64_t g = (n + threads - 1) / threads;
  return (unsigned)(g < 1 ? 1 : g);
}

// sum over the G lanes of a group, in every lane
template <int G, class T>
GSAP_D T group_allsum(T x) {
#pragma unroll
  for (int off = G / 2; off >= 1; off >>= 1) x += __shfl_xor(x, off, G);
  return x;
}

// First entry of p[0], p[stride], ..., p[(n - 1) stride] whose running weight exceeds
// target, in rounds of G x U entries: every lane sums its U loads (all in flight), the group
// total decides whether the crossing lies in this round, and only that round is scanned in
// order.  A zero entry never holds the crossing, so this equals the walk over the same row's
// nonzero entries in the CSR (or CSC) order.  fallback when the weights never exceed target.
template <int G, int U>
GSAP_D int32_t dense_walk(const int32_t *p, int64_t stride, int32_t n, uint64_t target,
                          int32_t fallback, uint64_t *before = nullptr) {
  const int q = threadIdx.x & (G - 1);
  const int gbase = (threadIdx.x & 63) & ~(G - 1);
  const uint64_t gm = G == 64 ? ~0ull : ((1ull << G) - 1ull);
  uint64_t base = 0;
  for (int32_t c = 0; c < n; c += G * U) {
    uint32_t w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t k = c + u * G + q;
      w[u] = k < n ? (uint32_t)p[(int64_t)k * stride] : 0u;
    }
    uint64_t s = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) s += w[u];
    const uint64_t tot = group_allsum<G>(s);
    if (base + tot > target) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        uint64_t incl = w[u];
#pragma unroll
        for (int off = 1; off < G; off <<= 1) {
          const uint64_t y = __shfl_up(incl, off, G);
          if (q >= off) incl += y;
        }
        const uint64_t bal = (__ballot(base + incl > target) >> gbase) & gm;
        if (bal) {
          const int first = __ffsll((long long)bal) - 1;
          // (before: the weight ahead of the entry found)
          if (before) *before = base + __shfl(incl - w[u], first, G);
          return c + u * G + first;
        }
        base += __shfl(incl, G - 1, G);
      }
    }
    base += tot;
  }
  return fallback;
}

// The same entry found from the end: with total = the weights' sum (> target), it is the
// last entry whose suffix sum (itself included) reaches total - target, so a target in the
// upper half of the mass is found reading the row's tail instead of its head.  Rounds of
// G x U entries from the end; inside the round holding the crossing, chunks from the last
// and, in a chunk, the highest lane whose suffix reaches the need.
template <int G, int U>
GSAP_D int32_t dense_walk_back(const int32_t *p, int64_t stride, int32_t n, uint64_t target,
                               uint64_t total, int32_t fallback) {
  const int q = threadIdx.x & (G - 1);
  const int gbase = (threadIdx.x & 63) & ~(G - 1);
  const uint64_t gm = G == 64 ? ~0ull : ((1ull << G) - 1ull);
  const uint64_t need = total - target;
  uint64_t base = 0;  // weight after the current round
  for (int32_t c = n - G * U; c > -G * U; c -= G * U) {
    uint32_t w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int32_t k = c + u * G + q;
      w[u] = k >= 0 ? (uint32_t)p[(int64_t)k * stride] : 0u;
    }
    uint64_t s = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) s += w[u];
    const uint64_t tot = group_allsum<G>(s);
    if (base + tot >= need) {
#pragma unroll
      for (int u = U - 1; u >= 0; --u) {
        uint64_t suf = w[u];  // this lane's weight and those of the higher lanes
#pragma unroll
        for (int off = 1; off < G; off <<= 1) {
          const uint64_t y = __shfl_down(suf, off, G);
          if (q + off < G) suf += y;
        }
        const uint64_t bal = (__ballot(base + suf >= need) >> gbase) & gm;
        if (bal) return c + u * G + (63 - __clzll((long long)bal));
        base += __shfl(suf, 0, G);
      }
    }
    base += tot;
  }
  return fallback;
}

// Group-parallel form of bm_walk (row x then column x; the first entry whose running weight
// exceeds target, skipping id skip with its mass carried forward): the target lies in the row
// when it is below the row's total d_out[x], else in the column at target - d_out[x]; that
// part's weights alone are walked (from the end nearer to the target in weight), which finds
// the crossing entry; the answer is the first entry from there whose id is not skip.  Same
// result as the sequential walk (every stored entry weighs at least 1).
template <int G>
GSAP_D int32_t group_walk(const BMView &m, int32_t x, uint64_t target, int32_t skip,
                          int32_t fallback) {
  const int32_t r0 = m.rowptr[x], nr = m.rowptr[x + 1] - r0;
  const int32_t c0 = m.colptr[x], nc = m.colptr[x + 1] - c0;
  const uint64_t dox = (uint64_t)m.dout[x], dix = (uint64_t)m.din[x];
  if (target >= dox + dix) return fallback;
  bool row = target < dox;
  const uint64_t t = row ? target : target - dox;
  const uint64_t tot = row ? dox : dix;
  const int32_t *w = row ? m.val + r0 : m.cval + c0;
  const int32_t n = row ? nr : nc;
  int32_t k = 2 * t < tot ? dense_walk<G, 4>(w, 1, n, t, 0)
                          : dense_walk_back<G, 4>(w, 1, n, t, tot, 0);
  while (true) {
    const int32_t id = row ? m.col[r0 + k] : m.row[c0 + k];
    if (id != skip) return id;
    ++k;
    if (row && k == nr) {
      row = false;
      k = 0;
    }
    if (!row && k >= nc) return fallback;
  }
}

// ------------------------------------------------ per-block terms and samples
// OUT[x] = sum_t M_xt (log2 M_xt - (log2 d_out_x + log2 d_in_t))  (kernels.hpp:143-182), used
// by the merge delta-S, and the nodal per-block neighbour sample and coin (kernels.hpp:942-997,
// 1024-1029; one draw per block per batch, shared by its vertices).
// (zero, when given: NCLS words cleared for the nodal proposal pass that follows; blocks
// whose row and column hold more than big entries are sampled by k_block_walk_big)
template <int G>
__global__ __launch_bounds__(256) void k_block_terms(BMView m, double *OUT, int32_t *sof,
                                                     uint8_t *coin, RngKey key, PhaseRef pr,
                                                     int sample, int32_t *zero, int32_t big) {
  if (zero && blockIdx.x == 0 && threadIdx.x < NCLS) zero[threadIdx.x] = 0;
  if (OUT) log2_tables_to_lds();  // only the row terms take logarithms
  const int32_t x = (int32_t)(blockIdx.x * (256 / G) + threadIdx.x / G);
  if (x >= m.B) return;
  const int q = threadIdx.x & (G - 1);
  if (OUT) {
    CAcc<G> acc;
    acc.init();
    const int32_t b0 = m.rowptr[x], n = m.rowptr[x + 1] - b0;
    const double lo = m.ldout[x];
    seg_for<G>(n, [&](int32_t k, int j) {
      acc.p[j] += term_s(m.val[b0 + k], lo, m.ldin[m.col[b0 + k]]);
    });
    const double out = acc.reduce();
    if (q == 0) OUT[x] = out;
  }
  if (sample) {
    const int32_t n = (m.rowptr[x + 1] - m.rowptr[x]) + (m.colptr[x + 1] - m.colptr[x]);
    if (n > big) return;
    const uint32_t phase = (pr.dev ? *pr.dev : 0u) + pr.add;
    const Philox y = philox((uint32_t)x, phase, key.chain, TAG_BLOCK, key.k0, key.k1);
    const int64_t deg = (int64_t)m.dout[x] + m.din[x];
    const int32_t so = deg == 0 ? 0 : group_walk<G>(m, x, mulhi(y.x0, deg), -1, 0);
    if (q == 0) {
      sof[x] = so;
      coin[x] = coin_random(y.x1, deg, m.B) ? 1 : 0;
    }
  }
}

// The neighbour sample of the blocks whose row and column hold more than big entries (the
// few blocks that greedy merges grow large), one 1024-thread workgroup per block: rounds of
// sixteen 1024-entry chunks, a chunk per wave summed with all its loads issued together,
// then a prefix over the chunk sums finds the chunk where the running weight crosses the
// target and its wave walks it.  Same entry as the sequential walk (integer sums).
constexpr int BT_BIG = 8192;     // entries of row + column above which a block walks here
constexpr int BT_BIG_MIN_B = 2048;  // smallest B with the separate walk
__global__ __launch_bounds__(1024) void k_block_walk_big(BMView m, int32_t *sof, uint8_t *coin,
                                                         RngKey key, PhaseRef pr, int32_t big) {
  __shared__ int32_t list[1024];
  __shared__ int32_t nlist;
  __shared__ unsigned long long wsum[16];
  __shared__ int32_t found;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t phase = (pr.dev ? *pr.dev : 0u) + pr.add;
  for (int64_t base = (int64_t)blockIdx.x * 1024; base < m.B; base += (int64_t)gridDim.x * 1024) {
    if (threadIdx.x == 0) nlist = 0;
    __syncthreads();
    const int64_t xt = base + threadIdx.x;
    if (xt < m.B) {
      const int32_t x = (int32_t)xt;
      const int32_t n = (m.rowptr[x + 1] - m.rowptr[x]) + (m.colptr[x + 1] - m.colptr[x]);
      if (n > big) list[atomicAdd(&nlist, 1)] = x;
    }
    __syncthreads();
    const int32_t nl = nlist;
    for (int32_t li = 0; li < nl; ++li) {
      const int32_t x = list[li];
      const Philox y = philox((uint32_t)x, phase, key.chain, TAG_BLOCK, key.k0, key.k1);
      const int64_t deg = (int64_t)m.dout[x] + m.din[x];
      const uint64_t target = mulhi(y.x0, deg);
      const int32_t r0 = m.rowptr[x], nr = m.rowptr[x + 1] - r0;
      const int32_t c0 = m.colptr[x], n = nr + (m.colptr[x + 1] - c0);
      if (threadIdx.x == 0) found = 0;
      __syncthreads();
      uint64_t rbase = 0;  // running weight before this round
      for (int32_t c0r = 0; c0r < n && deg > 0; c0r += 16 * 1024) {
        // this wave's chunk: entries c0r + w * 1024 + u * 64 + lane
        const int32_t k0 = c0r + w * 1024;
        uint32_t wt[16];
        int32_t id[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int32_t k = k0 + u * 64 + lane;
          wt[u] = 0;
          id[u] = 0;
          if (k < nr) { wt[u] = (uint32_t)m.val[r0 + k]; id[u] = m.col[r0 + k]; }
          else if (k < n) { wt[u] = (uint32_t)m.cval[c0 + k - nr]; id[u] = m.row[c0 + k - nr]; }
        }
        uint64_t s = 0;
#pragma unroll
        for (int u = 0; u < 16; ++u) s += wt[u];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) wsum[w] = s;
        __syncthreads();
        uint64_t pre = rbase, tot = 0;
        for (int i = 0; i < 16; ++i) {
          const uint64_t c = wsum[i];
          if (i < w) pre += c;
          tot += c;
        }
        // the chunk where the running weight first exceeds the target
        if (pre <= target && target < pre + s) {
          uint64_t b = pre;
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            uint64_t incl = wt[u];
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
              const uint64_t yv = __shfl_up(incl, off, 64);
              if (lane >= off) incl += yv;
            }
            const uint64_t bal = __ballot(b + incl > target);
            if (bal) {
              const int32_t pick = __shfl(id[u], __ffsll((long long)bal) - 1, 64);
              if (lane == 0) {
                sof[x] = pick;
                found = 1;
              }
              break;
            }
            b += __shfl(incl, 63, 64);
          }
        }
        __syncthreads();
        const int32_t done = found;
        rbase += tot;
        __syncthreads();  // wsum is rewritten by the next round
        if (done) break;
      }
      if (threadIdx.x == 0) {
        if (deg == 0) sof[x] = 0;
        coin[x] = coin_random(y.x1, deg, m.B) ? 1 : 0;
      }
      __syncthreads();
    }
    __syncthreads();  // the list is rewritten by the next range of blocks
  }
}

// The nodal per-block neighbour sample and coin (kernels.hpp:942-997, 1024-1029) while the
// CSR/CSC lags the labels (between the batches of a sweep at a dense B): the walk over row x
// and then column x of the dense matrix D, column x read as row x of its transpose DT when
// that is current (else down D's column).  The target lies in the row when it is below the
// row's total d_out[x], else in the column at target - d_out[x], so only that part is walked.
// Also writes the log2 block degrees of the degrees the batch update left in dout / din.
template <int G, int U>
__global__ __launch_bounds__(256) void k_block_terms_dense(const int32_t *D, const int32_t *DT,
                                                           const int32_t *CS,
                                                           int32_t B, const int32_t *dout,
                                                           const int32_t *din, double *ldout,
                                                           double *ldin, int32_t *sof,
                                                           uint8_t *coin, RngKey key,
                                                           PhaseRef pr, int32_t *zero) {
  if (zero && blockIdx.x == 0 && threadIdx.x < NCLS) zero[threadIdx.x] = 0;
  const int32_t x = (int32_t)(blockIdx.x * (256 / G) + threadIdx.x / G);
  if (x >= B) return;
  const int q = threadIdx.x & (G - 1);
  const int32_t dox = dout[x], dix = din[x];
  const uint32_t phase = (pr.dev ? *pr.dev : 0u) + pr.add;
  const Philox y = philox((uint32_t)x, phase, key.chain, TAG_BLOCK, key.k0, key.k1);
  const int64_t deg = (int64_t)dox + dix;
  int32_t so = 0;
  if (deg > 0) {
    const uint64_t target = mulhi(y.x0, deg);
    // the part holding the target (row x below d_out[x], else column x), walked from the
    // end nearer to it in weight
    const bool in_row = target < (uint64_t)dox;
    const uint64_t t = in_row ? target : target - (uint64_t)dox;
    const uint64_t tot = in_row ? (uint64_t)dox : (uint64_t)dix;
    if (!in_row && !DT && CS) {
      // column x through its chunk sums: the chunk holding the crossing, then its 64 rows
      const int32_t nch = (B + 63) / 64;
      uint64_t ahead = 0;
      const int32_t c = dense_walk<G, U>(CS + (int64_t)x * nch, 1, nch, t, 0, &ahead);
      const int32_t r0 = c * 64, nr = B - r0 < 64 ? B - r0 : 64;
      so = r0 + dense_walk<G, U>(D + (int64_t)r0 * B + x, B, nr, t - ahead, 0);
    } else {
      const int32_t *p = in_row ? D + (int64_t)x * B : DT ? DT + (int64_t)x * B : D + x;
      const int64_t stride = in_row || DT ? 1 : B;
      so = 2 * t < tot ? dense_walk<G, U>(p, stride, B, t, 0)
                       : dense_walk_back<G, U>(p, stride, B, t, tot, 0);
    }
  }
  if (q == 0) {
    sof[x] = so;
    coin[x] = coin_random(y.x1, deg, B) ? 1 : 0;
    ldout[x] = clog2((double)dox);
    ldin[x] = clog2((double)dix);
  }
}

void launch_block_terms_dense(const int32_t *D, const int32_t *DT, const int32_t *CS, int32_t B,
                              const int32_t *dout, const int32_t *din, double *ldout,
                              double *ldin, int32_t *sof, uint8_t *coin, RngKey key,
                              PhaseRef phase, hipStream_t st, int32_t *zero) {
  if (B <= 128)
    k_block_terms_dense<16, 4><<<grid_for(B, 16), 256, 0, st>>>(D, DT, CS, B, dout, din, ldout,
                                                               ldin, sof, coin, key, phase, zero);
  else if (B <= 1024)
    k_block_terms_dense<32, 8><<<grid_for(B, 8), 256, 0, st>>>(D, DT, CS, B, dout, din, ldout,
                                                              ldin, sof, coin, key, phase, zero);
  else
    k_block_terms_dense<64, 16><<<grid_for(B, 4), 256, 0, st>>>(D, DT, CS, B, dout, din, ldout,
                                                               ldin, sof, coin, key, phase, zero);
}

void launch_block_terms(const BMView &m, double *OUT, int32_t *sof, uint8_t *coin, RngKey key,
                        PhaseRef phase, bool sample, int32_t avg_row, hipStream_t st,
                        int32_t *zero) {
  const int s = sample ? 1 : 0;
  // the samples of very long rows and columns walk in a kernel of their own (B >= 2048)
  const bool sep = sample && m.B >= BT_BIG_MIN_B;
  const int32_t big = sep ? BT_BIG : INT32_MAX;
  if (avg_row <= 12)
    k_block_terms<8><<<grid_for(m.B, 32), 256, 0, st>>>(m, OUT, sof, coin, key, phase, s, zero,
                                                        big);
  else if (avg_row <= 24)
    k_block_terms<16><<<grid_for(m.B, 16), 256, 0, st>>>(m, OUT, sof, coin, key, phase, s, zero,
                                                         big);
  else if (avg_row <= 48)
    k_block_terms<32><<<grid_for(m.B, 8), 256, 0, st>>>(m, OUT, sof, coin, key, phase, s, zero,
                                                        big);
  else
    k_block_terms<64><<<grid_for(m.B, 4), 256, 0, st>>>(m, OUT, sof, coin, key, phase, s, zero,
                                                        big);
  if (sep) {
    const unsigned gb = (unsigned)std::min<int64_t>(1024, ((int64_t)m.B + 1023) / 1024);
    k_block_walk_big<<<gb, 1024, 0, st>>>(m, sof, coin, key, phase, big);
  }
}

// ------------------------------------------------------------- nodal moves
constexpr int LBITS = 8, LCAP = 1 << LBITS;

struct NodalArgs {
  int64_t lo, BS;
  const int32_t *optr, *oadj, *ow, *iptr, *iadj, *iw, *kout, *kin;
  LabView lab;
  BMView m;
  const int32_t *sof;
  const uint8_t *coin;
  RngKey key;
  PhaseRef phase;
  int32_t *prop;
  double *dS;
  const int64_t *hoff;
  const uint8_t *hbits_o, *hbits_i;
  int32_t *hpool;
  float *err;      // mixed form: |dS32 - dS64| where FP64 was evaluated, else 0
  uint8_t *near;   // mixed form: 0 FP32 rejection, 1 FP64 evaluated, 2 |dS32| within bound
  const int32_t *sloop;  // exact delta: self-loop weight per vertex
  double beta;           // Metropolis-Hastings inverse temperature
  int32_t pre;           // 1: prop already holds every proposal (k_nodal_propose)
  // with pre: the vertices whose proposal is another block, listed by class by the proposal
  // kernel (class c at dlist + c * BS, dcnt[c] of them) as records {v, optr[v], iptr[v],
  // out-degree | in-degree << 16} (degrees packed below the largest class, 0 there), and
  // the class bounds
  int4 *dlist;
  int32_t *dcnt;
  int32_t cb16, cb32, cbwave;
  // with pre, or null: running totals of the vertices listed for evaluation and of their
  // out+in adjacency entries, in EV_SLOTS pairs (summed by the host)
  unsigned long long *evstat;
};

// Variants of the nodal kernels (template bits): the exact delta-S and Metropolis-Hastings
// acceptance.
constexpr int VAR_EXACT = 1, VAR_MH = 2;

GSAP_D uint32_t nodal_phase(const NodalArgs &a) {
  return (a.phase.dev ? *a.phase.dev : 0u) + a.phase.add;
}

// A vertex's adjacency ranges: from its proposal record (unit weights, classes below the
// largest) or from the row pointers
struct VHead {
  int64_t v;
  int32_t o0, no, i0, ni;
};
GSAP_D VHead vhead_rec(const int4 R) {
  return VHead{(int64_t)R.x, R.y, R.w & 0xFFFF, R.z, (int32_t)((uint32_t)R.w >> 16)};
}
GSAP_D VHead vhead_load(const NodalArgs &a, int64_t v) {
  const int32_t o0 = a.optr[v], i0 = a.iptr[v];
  return VHead{v, o0, a.optr[v + 1] - o0, i0, a.iptr[v + 1] - i0};
}

constexpr float U32 = 5.9604645e-08f;  // 2^-24

// Final delta-S of one proposal (calculate_dS_overall_nodal, kernels.hpp:600-643): the
// row and column sums minus the untouched-entry terms, in this order.
GSAP_D double nodal_total(double drow, double dcol, double Y_r, double Y_s, double Y2_r,
                          double Y2_s) {
  double d = drow - Y_r;
  d = d - Y_s;
  d = d + dcol;
  d = d - Y2_r;
  d = d - Y2_s;
  return d;
}

// Mixed form, first pass: true when the FP32 delta-S proves a rejection (dS64 >= 0).
// eb is the summed error weight of the FP32 terms; 16 u eb bounds |dS32 - dS64| with
// margin (the FP32 term, slot and Kahan roundings need 11 u eb), plus the FP64
// roundings of the final combination.
GSAP_D bool mixed_rejects(double d32, float eb, double Y_r, double Y_s, double Y2_r,
                          double Y2_s, double &bound) {
  bound = 16.0 * (double)U32 * (double)eb +
          1.8e-15 * ((((__builtin_fabs(d32) + __builtin_fabs(Y_r)) + __builtin_fabs(Y_s)) +
                      __builtin_fabs(Y2_r)) + __builtin_fabs(Y2_s));
  return d32 >= bound;
}

// Metropolis-Hastings decision: accept iff u < exp(-beta dS) H, compared in log2, with
// u = (x + 1) 2^-32 from the vertex's third Philox word and lh = log2 H.
GSAP_D bool mh_accepts(double d, double lh, double beta, uint32_t x) {
  const double L = ((-beta) * d) * 1.4426950408889634 + lh;
  return clog2_s((double)x + 1.0) - 32.0 < L;
}

// Per-vertex constants of the Hastings ratio.  The nodal sampler proposes s for v with
// probability sum_t (k_t / k_v) (M_ts + M_st + 1) / (d_t + B) (neighbour block t, then a
// uniform block with probability B / (d_t + B), else a block of t's row and column weighted
// by M; d_t = d_out_t + d_in_t); the reverse proposal uses the state after the move.
struct Hast {
  int64_t v;
  int32_t r, s, B;
  int64_t Mrs, Msr, Mss;  // before the move
  int64_t Rr, Rs, Sr;     // M'_rr, M'_rs, M'_sr after the move
  int64_t kv, dr, ds;     // v's total degree; total degrees of r and s before the move
};

// True when the reverse term of an entry (neighbour u in block t) reads v's weights a_t, c_t.
GSAP_D bool hast_needs(const Hast &h, int32_t u, int32_t t) {
  return u != h.v && t != h.r && t != h.s;
}

// One adjacency entry's forward term w (M_ts + M_st + 1) / (d_t + B) and reverse term
// w (M'_t'r + M'_rt' + 1) / (d'_t' + B), t' = s for a self loop (u = v), else t; at, ct
// are v's out / in weights to block t (read only when hast_needs).
GSAP_D void hast_terms(const BMView &m, const Hast &h, int32_t u, int32_t t, int32_t w,
                       int64_t at, int64_t ct, double &fw, double &bk) {
  const int64_t Bq = h.B;
  int64_t nf, df, nb = 0, db = 1;
  if (t == h.r) {
    nf = h.Mrs + h.Msr + 1;
    df = h.dr + Bq;
  } else if (t == h.s) {
    nf = 2 * h.Mss + 1;
    df = h.ds + Bq;
  } else {
    int32_t Mrt, Mst, Mtr, Mts;
    bm_row_get2(m, h.r, h.s, t, Mrt, Mst);
    bm_col_get2(m, h.r, h.s, t, Mtr, Mts);
    nf = (int64_t)Mts + Mst + 1;
    df = (int64_t)m.dout[t] + m.din[t] + Bq;
    nb = ((int64_t)Mtr - ct) + ((int64_t)Mrt - at) + 1;
    db = df;
  }
  if (u == h.v || t == h.s) {
    nb = h.Sr + h.Rs + 1;
    db = h.ds + h.kv + Bq;
  } else if (t == h.r) {
    nb = 2 * h.Rr + 1;
    db = h.dr - h.kv + Bq;
  }
  fw = ((double)w * (double)nf) / (double)df;
  bk = ((double)w * (double)nb) / (double)db;
}

// log2 of the Hastings ratio from the canonical sums (0 for a vertex without edges)
GSAP_D double hast_log2(double fo, double fi, double bo, double bi, int32_t n) {
  return n == 0 ? 0.0 : clog2_s(bo + bi) - clog2_s(fo + fi);
}

// One wave per vertex v of the batch: proposal (kernels.hpp:877-939, 1000-1032) and delta-S
// against the batch-start block matrix (kernels.hpp:494-643, 1153-1726, node-only entries
// run-merged).  Only the entries v touches are evaluated term by term ("slots": v's
// out-blocks, r and s for rows r and s; v's in-blocks outside {r, s} for columns r and s);
// the untouched entries contribute (log2 d_old - log2 d_new) * (their summed count).
// Labels are not written here (batch-Jacobi).  Hash tables of 2^lbits slots live in LDS.

// unit: the head comes from a proposal record (unit weights: weights 1, degrees = counts)
template <bool MIX, int VAR>
GSAP_D void nodal_wave(const NodalArgs &a, const VHead hd, bool unit, int32_t *lds,
                       uint32_t lbits) {
  const int64_t v = hd.v;
  const uint32_t lcap = 1u << lbits;
  int32_t *lko = lds, *lwo = lds + lcap, *lfo = lds + 2 * lcap;
  int32_t *lki = lds + 3 * lcap, *lwi = lds + 4 * lcap, *lfi = lds + 5 * lcap;
  const int lane = threadIdx.x & 63;
  const int64_t i = v - a.lo;
  const BMView &m = a.m;
  const int32_t B = m.B;
  const int32_t r = a.lab[v];
  constexpr bool MHV = (VAR & VAR_MH) != 0;
  Philox x{0u, 0u, 0u, 0u};
  if (!a.pre || MHV)
    x = philox((uint32_t)v, nodal_phase(a), a.key.chain, TAG_NODE, a.key.k0, a.key.k1);
  const int32_t o0 = hd.o0, no = hd.no;
  const int32_t i0 = hd.i0, ni = hd.ni;
  // weights, degrees and block-matrix entries stay below 2^30 (the host rejects larger total
  // edge weights), so their sums and differences here fit int32 (fewer registers)
  const int32_t kout = unit ? no : a.kout[v], kin = unit ? ni : a.kin[v];
  const int32_t tot = kout + kin;
  // first chunk of both adjacency lists and its neighbour blocks, and the degrees of r:
  // gathered before the proposal, which they do not depend on, so that their loads overlap
  // its chain (when the proposal is r itself they go unused)
  const int32_t u_o0 = lane < no ? a.oadj[o0 + lane] : -1;
  const int32_t u_i0 = lane < ni ? a.iadj[i0 + lane] : -1;
  const int32_t w_o0 = lane < no ? (unit ? 1 : a.ow[o0 + lane]) : 0;
  const int32_t w_i0 = lane < ni ? (unit ? 1 : a.iw[i0 + lane]) : 0;
  const int32_t t_o0 = u_o0 >= 0 ? a.lab[u_o0] : -1;
  const int32_t t_i0 = u_i0 >= 0 ? a.lab[u_i0] : -1;
  const int32_t dout_r = m.dout[r], din_r = m.din[r];
  const double ldo_r = m.ldout[r], ldi_r = m.ldin[r];
  int32_t s;
  if (a.pre) {
    s = a.prop[i];
  } else if (tot == 0) {
    s = (int32_t)mulhi(x.x0, B);
  } else {
    const uint32_t target = mulhi(x.x1, tot);
    int32_t u;
    if (no <= 64 && ni <= 64) {
      // both lists inside the first chunk: the same pick as wave_pick (first entry of the
      // out list, then the in list, whose running weight exceeds target), and its block,
      // from the registers
      const uint32_t io = wave_incl_scan_u32((uint32_t)w_o0);
      const uint32_t ii = wave_incl_scan_u32((uint32_t)w_i0);
      const uint32_t tot_o = __shfl(io, 63, 64);
      const uint64_t mo = __ballot(lane < no && io > target);
      if (mo) {
        u = __shfl(t_o0, __ffsll((long long)mo) - 1, 64);
      } else {
        const uint64_t mi = __ballot(lane < ni && tot_o + ii > target);
        u = __shfl(t_i0, __ffsll((long long)mi) - 1, 64);
      }
    } else {
      const int32_t nbr =
          wave_pick(a.ow + o0, a.oadj + o0, no, a.iw + i0, a.iadj + i0, ni, target);
      u = a.lab[nbr];
    }
    s = a.coin[u] ? (int32_t)mulhi(x.x0, B) : a.sof[u];
  }
  if (lane == 0 && !a.pre) a.prop[i] = s;
  if (s == r) {
    if (lane == 0) {
      a.dS[i] = 0.0;
      if constexpr (MIX) {
        a.err[i] = 0.0f;
        a.near[i] = 0;
      }
    }
    return;
  }

  // a_t (out weight of v into block t) and c_t (in weight from block t), with the
  // first adjacency position of each block
  // (vertices here have at most lcap / 2 out- and in-neighbours)
  const HTab ho{lko, lwo, lfo, lbits};
  const HTab hi{lki, lwi, lfi, lbits};
  ht_clear(ho, lane, 64);
  ht_clear(hi, lane, 64);
  // block degrees of s, needed after the tables
  const int32_t dout_s = m.dout[s], din_s = m.din[s];
  const double ldo_s = m.ldout[s], ldi_s = m.ldin[s];
  wave_lds_sync();
  if (lane < no) ht_insert(ho, t_o0, w_o0, lane);
  if (lane < ni) ht_insert(hi, t_i0, w_i0, lane);
  for (int32_t j = lane + 64; j < no; j += 64) ht_insert(ho, a.lab[a.oadj[o0 + j]], a.ow[o0 + j], j);
  for (int32_t j = lane + 64; j < ni; j += 64) ht_insert(hi, a.lab[a.iadj[i0 + j]], a.iw[i0 + j], j);
  wave_lds_sync();
  const int32_t c_r = ht_weight(hi, r), c_s = ht_weight(hi, s);
  const bool r_in_a = ht_find(ho, r) >= 0, s_in_a = ht_find(ho, s) >= 0;
  constexpr bool EX = (VAR & VAR_EXACT) != 0, MH = (VAR & VAR_MH) != 0;
  const int32_t sig = EX ? a.sloop[v] : 0;  // a self loop moves from (r, r) to (s, s)
  // the Hastings ratio takes the move's real block matrix, self loop included, whatever
  // the delta-S form
  const int32_t sl = MH ? (EX ? sig : (int32_t)a.sloop[v]) : 0;
  const double l_dr = clog2_s((double)(dout_r - kout));
  const double l_ds = clog2_s((double)(dout_s + kout));
  const double l_ir = clog2_s((double)(din_r - kin));
  const double l_is = clog2_s((double)(din_s + kin));
  // M[r][r], M[s][r], M[r][s], M[s][s] for the untouched column counts, one lane each
  int32_t corner = 0;
  if (lane < 4) {
    const int32_t cc = lane < 2 ? r : s, rr = (lane & 1) ? s : r;
    corner = bm_col_get(m, cc, rr);
  }

  // rows r and s over the row slots; gathers shared by the FP32 and FP64 passes
  auto row_slot = [&](int32_t j, int32_t &t, int32_t &Mr, int32_t &Ms, int32_t &R,
                      int32_t &S) -> bool {
    int32_t av = 0;
    bool use;
    if (j < no) {
      t = j < 64 ? t_o0 : a.lab[a.oadj[o0 + j]];
      const int32_t slot = ht_find(ho, t);
      use = ld_relaxed(&ho.first[slot]) == j;
      av = ld_relaxed(&ho.w[slot]);
    } else if (j == no) {
      t = r;
      use = !r_in_a;
    } else {
      t = s;
      use = !s_in_a;
    }
    if (!use) return false;
    int32_t Mr32, Ms32;
    bm_row_get2(m, r, s, t, Mr32, Ms32);
    Mr = Mr32;
    Ms = Ms32;
    R = Mr - av - (t == r ? c_r - sig : 0) + (t == s ? c_r - sig : 0);
    S = Ms + av - (t == r ? c_s + sig : 0) + (t == s ? c_s + sig : 0);
    return true;
  };
  // columns r and s over the column slots (t not in {r, s})
  auto col_slot = [&](int32_t j, int32_t &t, int32_t &Mr, int32_t &Ms, int32_t &c) -> bool {
    t = j < 64 ? t_i0 : a.lab[a.iadj[i0 + j]];
    const int32_t slot = ht_find(hi, t);
    if (ld_relaxed(&hi.first[slot]) != j || t == r || t == s) return false;
    c = ld_relaxed(&hi.w[slot]);
    int32_t Mr32, Ms32;
    bm_col_get2(m, r, s, t, Mr32, Ms32);
    Mr = Mr32;
    Ms = Ms32;
    return true;
  };
  int32_t sum_r = 0, sum_s = 0, csum_r = 0, csum_s = 0;
  // FP64 terms (and, when sums is set, the integer sums of the touched entries)
  auto fp64_pass = [&](double &drow, double &dcol, bool sums) {
    double p_row = 0.0;
    for (int32_t j = lane; j < no + 2; j += 64) {
      int32_t t;
      int32_t Mr, Ms, R, S;
      if (!row_slot(j, t, Mr, Ms, R, S)) continue;
      const double ldt = m.ldin[t];
      // exact form: the new rows see the new column degrees of r and s
      const double ldn = !EX ? ldt : t == r ? l_ir : t == s ? l_is : ldt;
      const double or_ = Mr > 0 ? term_s(Mr, ldo_r, ldt) : 0.0;
      const double os_ = Ms > 0 ? term_s(Ms, ldo_s, ldt) : 0.0;
      const double nr = R > 0 ? term_s(R, ldn, l_dr) : 0.0;
      const double ns = S > 0 ? term_s(S, ldn, l_ds) : 0.0;
      p_row += ((or_ + os_) - nr) - ns;
      if (sums) {
        sum_r += Mr;
        sum_s += Ms;
      }
    }
    double p_col = 0.0;
    for (int32_t j = lane; j < ni; j += 64) {
      int32_t t;
      int32_t Mr, Ms, c;
      if (!col_slot(j, t, Mr, Ms, c)) continue;
      const double ldt = m.ldout[t];
      const double ir = Mr > 0 ? term_s(Mr, ldt, ldi_r) : 0.0;
      const double is = Ms > 0 ? term_s(Ms, ldt, ldi_s) : 0.0;
      const int32_t wr = Mr - c, ws = Ms + c;
      const double nr = wr > 0 ? term_s(wr, ldt, l_ir) : 0.0;
      const double ns = ws > 0 ? term_s(ws, ldt, l_is) : 0.0;
      p_col += ((ir + is) - nr) - ns;
      if (sums) {
        csum_r += Mr;
        csum_s += Ms;
      }
    }
    drow = wave_sum(p_row);
    dcol = wave_sum(p_col);
  };
  // plain form: one FP64 pass; mixed form: an FP32 pass first (with the integer sums)
  double drow = 0.0, dcol = 0.0;
  Kahan krow, kcol;
  krow.init();
  kcol.init();
  float eb = 0.0f;
  if constexpr (!MIX) {
    fp64_pass(drow, dcol, true);
  } else {
    for (int32_t j = lane; j < no + 2; j += 64) {
      int32_t t;
      int32_t Mr, Ms, R, S;
      if (!row_slot(j, t, Mr, Ms, R, S)) continue;
      sum_r += Mr;
      sum_s += Ms;
      const float ldt = (float)m.ldin[t];
      const float ldn = !EX ? ldt : t == r ? (float)l_ir : t == s ? (float)l_is : ldt;
      const float or_ = Mr > 0 ? term32(Mr, (float)ldo_r, ldt, eb) : 0.0f;
      const float os_ = Ms > 0 ? term32(Ms, (float)ldo_s, ldt, eb) : 0.0f;
      const float nr = R > 0 ? term32(R, ldn, (float)l_dr, eb) : 0.0f;
      const float ns = S > 0 ? term32(S, ldn, (float)l_ds, eb) : 0.0f;
      krow.add(((or_ + os_) - nr) - ns);
    }
    for (int32_t j = lane; j < ni; j += 64) {
      int32_t t;
      int32_t Mr, Ms, c;
      if (!col_slot(j, t, Mr, Ms, c)) continue;
      csum_r += Mr;
      csum_s += Ms;
      const float ldt = (float)m.ldout[t];
      const float ir = Mr > 0 ? term32(Mr, ldt, (float)ldi_r, eb) : 0.0f;
      const float is = Ms > 0 ? term32(Ms, ldt, (float)ldi_s, eb) : 0.0f;
      const int32_t wr = Mr - c, ws = Ms + c;
      const float nr = wr > 0 ? term32(wr, ldt, (float)l_ir, eb) : 0.0f;
      const float ns = ws > 0 ? term32(ws, ldt, (float)l_is, eb) : 0.0f;
      kcol.add(((ir + is) - nr) - ns);
    }
  }
  sum_r = wave_sum_i64(sum_r);
  sum_s = wave_sum_i64(sum_s);
  csum_r = wave_sum_i64(csum_r);
  csum_s = wave_sum_i64(csum_s);
  const int32_t M_rr = __shfl(corner, 0, 64), M_sr = __shfl(corner, 1, 64);
  const int32_t M_rs = __shfl(corner, 2, 64), M_ss = __shfl(corner, 3, 64);
  const int32_t W_r = dout_r - sum_r, W_s = dout_s - sum_s;
  const double Y_r = W_r > 0 ? (ldo_r - l_dr) * (double)W_r : 0.0;
  const double Y_s = W_s > 0 ? (ldo_s - l_ds) * (double)W_s : 0.0;
  const int32_t W2_r = din_r - M_rr - M_sr - csum_r;
  const int32_t W2_s = din_s - M_rs - M_ss - csum_s;
  const double Y2_r = W2_r > 0 ? (ldi_r - l_ir) * (double)W2_r : 0.0;
  const double Y2_s = W2_s > 0 ? (ldi_s - l_is) * (double)W2_s : 0.0;
  double d32 = 0.0, bound = 0.0;
  if constexpr (MIX) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      krow.combine(__shfl_down(krow.s, off, 64), __shfl_down(krow.c, off, 64));
      kcol.combine(__shfl_down(kcol.s, off, 64), __shfl_down(kcol.c, off, 64));
      eb += __shfl_down(eb, off, 64);
    }
    int rej = 0;
    if (lane == 0) {
      d32 = nodal_total((double)krow.value(), (double)kcol.value(), Y_r, Y_s, Y2_r, Y2_s);
      rej = mixed_rejects(d32, eb, Y_r, Y_s, Y2_r, Y2_s, bound);
    }
    if (__shfl(rej, 0, 64)) {
      if (lane == 0) {
        a.dS[i] = 0.0;
        a.err[i] = 0.0f;
        a.near[i] = 0;
      }
      return;
    }
    fp64_pass(drow, dcol, false);
  }
  double lh = 0.0;
  if constexpr (MH) {
    const int32_t a_r = ht_weight(ho, r), a_s = ht_weight(ho, s);
    const Hast h{v, r, s, B, M_rs, M_sr, M_ss, M_rr - a_r - c_r + sl, M_rs - a_s + c_r - sl,
                 M_sr + a_r - c_s - sl, tot, dout_r + din_r, dout_s + din_s};
    double fo = 0.0, fi = 0.0, bo = 0.0, bi = 0.0;
    for (int32_t j = lane; j < no; j += 64) {
      const int32_t u = a.oadj[o0 + j];
      const int32_t t = j < 64 ? t_o0 : a.lab[u];
      const bool nd = hast_needs(h, u, t);
      double f, b;
      hast_terms(m, h, u, t, j < 64 ? w_o0 : a.ow[o0 + j], nd ? ht_weight(ho, t) : 0,
                 nd ? ht_weight(hi, t) : 0, f, b);
      fo += f;
      bo += b;
    }
    for (int32_t j = lane; j < ni; j += 64) {
      const int32_t u = a.iadj[i0 + j];
      const int32_t t = j < 64 ? t_i0 : a.lab[u];
      const bool nd = hast_needs(h, u, t);
      double f, b;
      hast_terms(m, h, u, t, j < 64 ? w_i0 : a.iw[i0 + j], nd ? ht_weight(ho, t) : 0,
                 nd ? ht_weight(hi, t) : 0, f, b);
      fi += f;
      bi += b;
    }
    fo = wave_sum(fo);
    fi = wave_sum(fi);
    bo = wave_sum(bo);
    bi = wave_sum(bi);
    if (lane == 0) lh = hast_log2(fo, fi, bo, bi, no + ni);
  }
  if (lane == 0) {
    const double d = nodal_total(drow, dcol, Y_r, Y_s, Y2_r, Y2_s);
    if constexpr (MH) {
      const bool acc = mh_accepts(d, lh, a.beta, x.x2);
      a.prop[i] = acc ? s : r;
      a.dS[i] = acc ? d : 0.0;
      return;
    }
    a.dS[i] = d < 0.0 ? d : 0.0;  // accept iff dS < 0
    if constexpr (MIX) {
      // non-finite delta-S (an empty block's -inf log, the reference's quirk) only ever
      // sends a proposal to the FP64 pass; it is left out of the error statistic
      const double e = __builtin_fabs(d32 - d);
      a.err[i] = e <= 1e30 ? (float)e : 0.0f;
      a.near[i] = d32 > -bound ? 2 : 1;
    }
  }
}

// One 256-thread workgroup per vertex with more than 128 out- or in-neighbours: the
// computation of k_nodal with the adjacency spread over four waves.  Hash tables are sized
// for the vertex (in LDS when they fit the launch's allocation, else in the vertex's slice
// of the global pool); every row and column slot value is stored by position and wave 0
// sums them in the canonical order, so the bits equal the one-wave form.  The mixed form
// evaluates these few vertices in FP64 directly.
template <bool MIX, int VAR>
GSAP_D void nodal_big(const NodalArgs &a, int64_t v, double *lds_d, uint32_t lds_bytes) {
  __shared__ int64_t red[4][4];
  __shared__ int64_t corner_s[4];
  __shared__ uint32_t pick_sc[8];
  __shared__ double dh_s[2];  // Metropolis-Hastings: delta-S and the forward sum
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t i = v - a.lo;
  const BMView &m = a.m;
  const int32_t B = m.B;
  const int32_t r = a.lab[v];
  constexpr bool MHV = (VAR & VAR_MH) != 0;
  Philox x{0u, 0u, 0u, 0u};
  if (!a.pre || MHV)
    x = philox((uint32_t)v, nodal_phase(a), a.key.chain, TAG_NODE, a.key.k0, a.key.k1);
  const int32_t o0 = a.optr[v], no = a.optr[v + 1] - o0;
  const int32_t i0 = a.iptr[v], ni = a.iptr[v + 1] - i0;
  const int64_t kout = a.kout[v], kin = a.kin[v];
  const int64_t tot = kout + kin;
  int32_t s;  // every thread computes the same proposal
  if (a.pre) {
    s = a.prop[i];
  } else if (tot == 0) {
    s = (int32_t)mulhi(x.x0, B);
  } else {
    const int32_t nbr = block_pick(a.ow + o0, a.oadj + o0, no, a.iw + i0, a.iadj + i0, ni,
                                   mulhi(x.x1, tot), pick_sc);
    const int32_t u = a.lab[nbr];
    s = a.coin[u] ? (int32_t)mulhi(x.x0, B) : a.sof[u];
  }
  if (tid == 0 && !a.pre) a.prop[i] = s;
  if (s == r) {
    if (tid == 0) {
      a.dS[i] = 0.0;
      if constexpr (MIX) {
        a.err[i] = 0.0f;
        a.near[i] = 0;
      }
    }
    return;
  }
  const uint32_t bo = a.hbits_o[v], bi = a.hbits_i[v];
  const int32_t co = 1 << bo, ci = 1 << bi;
  const int64_t words = 3 * (int64_t)(co + ci) + 2 * (int64_t)(no + 2 + ni);
  int32_t *base = 4 * words <= (int64_t)lds_bytes ? reinterpret_cast<int32_t *>(lds_d)
                                                  : a.hpool + a.hoff[v];
  const HTab ho{base, base + co, base + 2 * co, bo};
  const HTab hi{base + 3 * co, base + 3 * co + ci, base + 3 * co + 2 * ci, bi};
  double *xs = reinterpret_cast<double *>(base + 3 * (co + ci));
  ht_clear(ho, tid, 256);
  ht_clear(hi, tid, 256);
  const int64_t dout_r = m.dout[r], dout_s = m.dout[s], din_r = m.din[r], din_s = m.din[s];
  const double ldo_r = m.ldout[r], ldo_s = m.ldout[s], ldi_r = m.ldin[r], ldi_s = m.ldin[s];
  if (tid < 4) {
    const int32_t cc = tid < 2 ? r : s, rr = (tid & 1) ? s : r;
    corner_s[tid] = bm_col_get(m, cc, rr);
  }
  __syncthreads();
  // four entries per thread in flight: ids and weights, then their blocks, then inserts
  auto insert_all = [&](const HTab &h, const int32_t *adj, const int32_t *wt, int32_t n) {
    for (int32_t j0 = tid; j0 < n; j0 += 4 * 256) {
      int32_t t[4], w[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int32_t j = j0 + q * 256;
        t[q] = j < n ? adj[j] : 0;
        w[q] = j < n ? wt[j] : 0;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (j0 + q * 256 < n) t[q] = a.lab[t[q]];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (j0 + q * 256 < n) ht_insert(h, t[q], w[q], j0 + q * 256);
    }
  };
  insert_all(ho, a.oadj + o0, a.ow + o0, no);
  insert_all(hi, a.iadj + i0, a.iw + i0, ni);
  __syncthreads();
  const int64_t c_r = ht_weight(hi, r), c_s = ht_weight(hi, s);
  const bool r_in_a = ht_find(ho, r) >= 0, s_in_a = ht_find(ho, s) >= 0;
  constexpr bool EX = (VAR & VAR_EXACT) != 0, MH = (VAR & VAR_MH) != 0;
  const int64_t sig = EX ? a.sloop[v] : 0;  // a self loop moves from (r, r) to (s, s)
  // the Hastings ratio takes the move's real block matrix, self loop included, whatever
  // the delta-S form
  const int64_t sl = MH ? (EX ? sig : (int64_t)a.sloop[v]) : 0;
  const double l_dr = clog2_s((double)(dout_r - kout));
  const double l_ds = clog2_s((double)(dout_s + kout));
  const double l_ir = clog2_s((double)(din_r - kin));
  const double l_is = clog2_s((double)(din_s + kin));
  int64_t sum_r = 0, sum_s = 0, csum_r = 0, csum_s = 0;
  // row slots: positions 0..no-1 (first occurrences), no (r), no+1 (s)
  for (int32_t j = tid; j < no + 2; j += 256) {
    int32_t t;
    int64_t av = 0;
    bool use;
    if (j < no) {
      t = a.lab[a.oadj[o0 + j]];
      const int32_t slot = ht_find(ho, t);
      use = ld_relaxed(&ho.first[slot]) == j;
      av = ld_relaxed(&ho.w[slot]);
    } else if (j == no) {
      t = r;
      use = !r_in_a;
    } else {
      t = s;
      use = !s_in_a;
    }
    double xv = 0.0;
    if (use) {
      int32_t Mr32, Ms32;
      bm_row_get2(m, r, s, t, Mr32, Ms32);
      const int64_t Mr = Mr32, Ms = Ms32;
      const int64_t R = Mr - av - (t == r ? c_r - sig : 0) + (t == s ? c_r - sig : 0);
      const int64_t S = Ms + av - (t == r ? c_s + sig : 0) + (t == s ? c_s + sig : 0);
      const double ldt = m.ldin[t];
      const double ldn = !EX ? ldt : t == r ? l_ir : t == s ? l_is : ldt;
      const double or_ = Mr > 0 ? term_s(Mr, ldo_r, ldt) : 0.0;
      const double os_ = Ms > 0 ? term_s(Ms, ldo_s, ldt) : 0.0;
      const double nr = R > 0 ? term_s(R, ldn, l_dr) : 0.0;
      const double ns = S > 0 ? term_s(S, ldn, l_ds) : 0.0;
      xv = ((or_ + os_) - nr) - ns;
      sum_r += Mr;
      sum_s += Ms;
    }
    xs[j] = xv;
  }
  // column slots (t not in {r, s})
  for (int32_t j = tid; j < ni; j += 256) {
    const int32_t t = a.lab[a.iadj[i0 + j]];
    const int32_t slot = ht_find(hi, t);
    double xv = 0.0;
    if (ld_relaxed(&hi.first[slot]) == j && t != r && t != s) {
      const int64_t c = ld_relaxed(&hi.w[slot]);
      int32_t Mr32, Ms32;
      bm_col_get2(m, r, s, t, Mr32, Ms32);
      const int64_t Mr = Mr32, Ms = Ms32;
      const double ldt = m.ldout[t];
      const double ir = Mr > 0 ? term_s(Mr, ldt, ldi_r) : 0.0;
      const double is = Ms > 0 ? term_s(Ms, ldt, ldi_s) : 0.0;
      const int64_t wr = Mr - c, ws = Ms + c;
      const double nr = wr > 0 ? term_s(wr, ldt, l_ir) : 0.0;
      const double ns = ws > 0 ? term_s(ws, ldt, l_is) : 0.0;
      xv = ((ir + is) - nr) - ns;
      csum_r += Mr;
      csum_s += Ms;
    }
    xs[no + 2 + j] = xv;
  }
  sum_r = wave_sum_i64(sum_r);
  sum_s = wave_sum_i64(sum_s);
  csum_r = wave_sum_i64(csum_r);
  csum_s = wave_sum_i64(csum_s);
  if (lane == 0) {
    red[wv][0] = sum_r;
    red[wv][1] = sum_s;
    red[wv][2] = csum_r;
    red[wv][3] = csum_s;
  }
  __syncthreads();
  if (!MH && wv != 0) return;
  if (wv == 0) {
  // canonical sums: position k into partial k mod 64, ascending; unused slots hold +0.0
  double p_row = 0.0, p_col = 0.0;
  for (int32_t k = lane; k < no + 2; k += 64) p_row += xs[k];
  for (int32_t k = lane; k < ni; k += 64) p_col += xs[no + 2 + k];
  const double drow = wave_sum(p_row), dcol = wave_sum(p_col);
  if (lane == 0) {
    const int64_t Sr = red[0][0] + red[1][0] + red[2][0] + red[3][0];
    const int64_t Ss = red[0][1] + red[1][1] + red[2][1] + red[3][1];
    const int64_t Cr = red[0][2] + red[1][2] + red[2][2] + red[3][2];
    const int64_t Cs = red[0][3] + red[1][3] + red[2][3] + red[3][3];
    const int64_t M_rr = corner_s[0], M_sr = corner_s[1], M_rs = corner_s[2], M_ss = corner_s[3];
    const int64_t W_r = dout_r - Sr, W_s = dout_s - Ss;
    const double Y_r = W_r > 0 ? (ldo_r - l_dr) * (double)W_r : 0.0;
    const double Y_s = W_s > 0 ? (ldo_s - l_ds) * (double)W_s : 0.0;
    const int64_t W2_r = din_r - M_rr - M_sr - Cr;
    const int64_t W2_s = din_s - M_rs - M_ss - Cs;
    const double Y2_r = W2_r > 0 ? (ldi_r - l_ir) * (double)W2_r : 0.0;
    const double Y2_s = W2_s > 0 ? (ldi_s - l_is) * (double)W2_s : 0.0;
    const double d = nodal_total(drow, dcol, Y_r, Y_s, Y2_r, Y2_s);
    if constexpr (MH) {
      dh_s[0] = d;
    } else {
      a.dS[i] = d < 0.0 ? d : 0.0;
      if constexpr (MIX) {
        a.err[i] = 0.0f;
        a.near[i] = 1;
      }
    }
  }
  }
  if constexpr (MH) {
    // Hastings sums in two passes over the slot array: forward terms, then reverse terms,
    // each summed canonically by wave 0 (out entries at 0.., in entries at no + 2 ..)
    const int64_t M_rr = corner_s[0], M_sr = corner_s[1], M_rs = corner_s[2], M_ss = corner_s[3];
    const int64_t a_r = ht_weight(ho, r), a_s = ht_weight(ho, s);
    const Hast h{v, r, s, B, M_rs, M_sr, M_ss, M_rr - a_r - c_r + sl, M_rs - a_s + c_r - sl,
                 M_sr + a_r - c_s - sl, tot, dout_r + din_r, dout_s + din_s};
    auto fill = [&](bool back) {
      for (int32_t j = tid; j < no + ni; j += 256) {
        const bool out = j < no;
        const int32_t k = out ? j : j - no;
        const int32_t u = out ? a.oadj[o0 + k] : a.iadj[i0 + k];
        const int32_t t = a.lab[u];
        const bool nd = back && hast_needs(h, u, t);
        double f, b;
        hast_terms(m, h, u, t, out ? a.ow[o0 + k] : a.iw[i0 + k], nd ? ht_weight(ho, t) : 0,
                   nd ? ht_weight(hi, t) : 0, f, b);
        xs[out ? k : no + 2 + k] = back ? b : f;
      }
    };
    auto sums = [&](double &so, double &si) {
      double po = 0.0, pi = 0.0;
      for (int32_t k = lane; k < no; k += 64) po += xs[k];
      for (int32_t k = lane; k < ni; k += 64) pi += xs[no + 2 + k];
      so = wave_sum(po);
      si = wave_sum(pi);
    };
    __syncthreads();  // wave 0 has read the delta-S slots
    fill(false);
    __syncthreads();
    if (wv == 0) {
      double fo, fi;
      sums(fo, fi);
      if (lane == 0) dh_s[1] = fo + fi;
    }
    __syncthreads();
    fill(true);
    __syncthreads();
    if (wv == 0) {
      double bo, bi;
      sums(bo, bi);
      if (lane == 0) {
        // same grouping as hast_log2: (bo + bi) and the forward (fo + fi)
        const double lh = no + ni == 0 ? 0.0 : clog2_s(bo + bi) - clog2_s(dh_s[1]);
        const double d = dh_s[0];
        const bool acc = mh_accepts(d, lh, a.beta, x.x2);
        a.prop[i] = acc ? s : r;
        a.dS[i] = acc ? d : 0.0;
      }
    }
  }
}

// Group sums over the G lanes of a group (every lane gets the result).
// Same computation as k_nodal for a vertex with at most G out- and G in-neighbours, G lanes
// per vertex: lane q holds out-entry q and in-entry q; the neighbour blocks are deduplicated
// in registers (shuffles inside the group) instead of hash tables, and the block of the
// sampled neighbour comes from the same gathered values.
// (rec: the proposal kernel's records, unit weights; else list)
template <int G, bool MIX, int VAR>
GSAP_D void nodal_small(const NodalArgs &a, const int32_t *list, const int4 *rec, int32_t n,
                        int32_t bid) {
  const int32_t gid = (int32_t)(bid * (256 / G) + threadIdx.x / G);
  if (gid >= n) return;
  const int q = threadIdx.x & (G - 1);
  const int gbase = (threadIdx.x & 63) & ~(G - 1);
  const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
  const bool unit = rec != nullptr;
  const VHead hd = unit ? vhead_rec(rec[gid]) : vhead_load(a, list[gid]);
  const int64_t v = hd.v;
  const int64_t i = v - a.lo;
  const BMView &m = a.m;
  const int32_t B = m.B;
  const int32_t r = a.lab[v];
  constexpr bool EX = (VAR & VAR_EXACT) != 0, MH = (VAR & VAR_MH) != 0;
  Philox x{0u, 0u, 0u, 0u};
  if (!a.pre || MH)
    x = philox((uint32_t)v, nodal_phase(a), a.key.chain, TAG_NODE, a.key.k0, a.key.k1);
  const int32_t o0 = hd.o0, no = hd.no;
  const int32_t i0 = hd.i0, ni = hd.ni;
  const int32_t kout = unit ? no : a.kout[v], kin = unit ? ni : a.kin[v];
  const int32_t tot = kout + kin;
  int32_t t_o = -1, w_o = 0, t_i = -1, w_i = 0, u_o = -1, u_i = -1;
  if (q < no) { w_o = unit ? 1 : a.ow[o0 + q]; u_o = a.oadj[o0 + q]; t_o = a.lab[u_o]; }
  if (q < ni) { w_i = unit ? 1 : a.iw[i0 + q]; u_i = a.iadj[i0 + q]; t_i = a.lab[u_i]; }
  // proposal: first entry of (out list, in list) whose running weight exceeds the target
  int32_t s;
  if (a.pre) {
    s = a.prop[i];
  } else if (tot == 0) {
    s = (int32_t)mulhi(x.x0, B);
  } else {
    const uint32_t target = mulhi(x.x1, tot);
    uint32_t io = (uint32_t)w_o, ii = (uint32_t)w_i;
#pragma unroll
    for (int off = 1; off < G; off <<= 1) {
      const uint32_t yo = __shfl_up(io, off, G), yi = __shfl_up(ii, off, G);
      if (q >= off) { io += yo; ii += yi; }
    }
    const uint32_t tot_o = __shfl(io, G - 1, G);
    const uint64_t mo = (__ballot(q < no && io > target) >> gbase) & gmask;
    int32_t u;
    if (mo) {
      u = __shfl(t_o, __ffsll((long long)mo) - 1, G);
    } else {
      const uint64_t mi = (__ballot(q < ni && tot_o + ii > target) >> gbase) & gmask;
      u = __shfl(t_i, __ffsll((long long)mi) - 1, G);
    }
    s = a.coin[u] ? (int32_t)mulhi(x.x0, B) : a.sof[u];
  }
  if (q == 0 && !a.pre) a.prop[i] = s;
  if (s == r) {
    if (q == 0) {
      a.dS[i] = 0.0;
      if constexpr (MIX) {
        a.err[i] = 0.0f;
        a.near[i] = 0;
      }
    }
    return;
  }
  // per-lane block weight and first-occurrence flag (Metropolis-Hastings also needs the
  // in weight of the lane's out block and the out weight of its in block)
  int32_t a_sum = 0, c_sum = 0, c_of_o = 0, a_of_i = 0;
  bool first_o = q < no, first_i = q < ni;
  const int pend = no > ni ? no : ni;
#pragma unroll 1
  for (int p = 0; p < pend; ++p) {
    const int32_t tpo = __shfl(t_o, p, G), wpo = __shfl(w_o, p, G);
    const int32_t tpi = __shfl(t_i, p, G), wpi = __shfl(w_i, p, G);
    if (tpo == t_o) { a_sum += wpo; if (p < q) first_o = false; }
    if (tpi == t_i) { c_sum += wpi; if (p < q) first_i = false; }
    if constexpr (MH) {
      if (tpi == t_o) c_of_o += wpi;
      if (tpo == t_i) a_of_i += wpo;
    }
  }
  const int32_t sig = EX ? a.sloop[v] : 0;  // a self loop moves from (r, r) to (s, s)
  // the Hastings ratio takes the move's real block matrix, self loop included, whatever
  // the delta-S form
  const int32_t sl = MH ? (EX ? sig : (int32_t)a.sloop[v]) : 0;
  const int32_t c_r = group_allsum<G>((q < ni && t_i == r) ? w_i : 0);
  const int32_t c_s = group_allsum<G>((q < ni && t_i == s) ? w_i : 0);
  const bool r_in_a = ((__ballot(q < no && t_o == r) >> gbase) & gmask) != 0;
  const bool s_in_a = ((__ballot(q < no && t_o == s) >> gbase) & gmask) != 0;
  const int32_t dout_r = m.dout[r], dout_s = m.dout[s], din_r = m.din[r], din_s = m.din[s];
  const double ldo_r = m.ldout[r], ldo_s = m.ldout[s], ldi_r = m.ldin[r], ldi_s = m.ldin[s];
  const double l_dr = clog2_s((double)(dout_r - kout));
  const double l_ds = clog2_s((double)(dout_s + kout));
  const double l_ir = clog2_s((double)(din_r - kin));
  const double l_is = clog2_s((double)(din_s + kin));
  int32_t corner = 0;
  if (q < 4) {
    const int32_t cc = q < 2 ? r : s, rr = (q & 1) ? s : r;
    corner = bm_col_get(m, cc, rr);
  }
  // rows r and s over the row slots: positions 0..no-1 (first occurrences), no (r), no+1 (s)
  auto row_slot = [&](int32_t j, int32_t &t, int32_t &Mr, int32_t &Ms, int32_t &R,
                      int32_t &S) -> bool {
    int32_t av = 0;
    bool use;
    if (j < no) {
      t = t_o;
      av = a_sum;
      use = first_o;
    } else if (j == no) {
      t = r;
      use = !r_in_a;
    } else {
      t = s;
      use = !s_in_a;
    }
    if (!use) return false;
    int32_t Mr32, Ms32;
    bm_row_get2(m, r, s, t, Mr32, Ms32);
    Mr = Mr32;
    Ms = Ms32;
    R = Mr - av - (t == r ? c_r - sig : 0) + (t == s ? c_r - sig : 0);
    S = Ms + av - (t == r ? c_s + sig : 0) + (t == s ? c_s + sig : 0);
    return true;
  };
  // column slot of lane q (t not in {r, s})
  const bool col_use = first_i && t_i != r && t_i != s;
  auto col_get = [&](int32_t &Mr, int32_t &Ms) {
    int32_t Mr32, Ms32;
    bm_col_get2(m, r, s, t_i, Mr32, Ms32);
    Mr = Mr32;
    Ms = Ms32;
  };
  int32_t sum_r = 0, sum_s = 0, csum_r = 0, csum_s = 0;
  // positions j = q + jj G (jj < JJ cover 0 .. no + 1 <= G + 1); position j belongs to
  // canonical partial j mod 64, which is this lane's p[jj mod (64 / G)]
  constexpr int JJ = G == 64 ? 2 : 64 / G;
  auto fp64_pass = [&](double &drow, double &dcol, bool sums) {
    CAcc<G> prow;
    prow.init();
#pragma unroll
    for (int jj = 0; jj < JJ; ++jj) {
      const int32_t j = q + jj * G;
      if (j >= no + 2) break;
      int32_t t;
      int32_t Mr, Ms, R, S;
      if (!row_slot(j, t, Mr, Ms, R, S)) continue;
      const double ldt = m.ldin[t];
      const double ldn = !EX ? ldt : t == r ? l_ir : t == s ? l_is : ldt;
      const double or_ = Mr > 0 ? term_s(Mr, ldo_r, ldt) : 0.0;
      const double os_ = Ms > 0 ? term_s(Ms, ldo_s, ldt) : 0.0;
      const double nr = R > 0 ? term_s(R, ldn, l_dr) : 0.0;
      const double ns = S > 0 ? term_s(S, ldn, l_ds) : 0.0;
      prow.p[jj % CAcc<G>::S] += ((or_ + os_) - nr) - ns;
      if (sums) {
        sum_r += Mr;
        sum_s += Ms;
      }
    }
    CAcc<G> pcol;
    pcol.init();
    if (col_use) {
      int32_t Mr, Ms;
      col_get(Mr, Ms);
      const double ldt = m.ldout[t_i];
      const double ir = Mr > 0 ? term_s(Mr, ldt, ldi_r) : 0.0;
      const double is = Ms > 0 ? term_s(Ms, ldt, ldi_s) : 0.0;
      const int32_t wr = Mr - c_sum, ws = Ms + c_sum;
      const double nr = wr > 0 ? term_s(wr, ldt, l_ir) : 0.0;
      const double ns = ws > 0 ? term_s(ws, ldt, l_is) : 0.0;
      pcol.p[0] += ((ir + is) - nr) - ns;
      if (sums) {
        csum_r += Mr;
        csum_s += Ms;
      }
    }
    drow = prow.reduce();
    dcol = pcol.reduce();
  };
  double drow = 0.0, dcol = 0.0;
  Kahan krow, kcol;
  krow.init();
  kcol.init();
  float eb = 0.0f;
  if constexpr (!MIX) {
    fp64_pass(drow, dcol, true);
  } else {
#pragma unroll
    for (int jj = 0; jj < JJ; ++jj) {
      const int32_t j = q + jj * G;
      if (j >= no + 2) break;
      int32_t t;
      int32_t Mr, Ms, R, S;
      if (!row_slot(j, t, Mr, Ms, R, S)) continue;
      sum_r += Mr;
      sum_s += Ms;
      const float ldt = (float)m.ldin[t];
      const float ldn = !EX ? ldt : t == r ? (float)l_ir : t == s ? (float)l_is : ldt;
      const float or_ = Mr > 0 ? term32(Mr, (float)ldo_r, ldt, eb) : 0.0f;
      const float os_ = Ms > 0 ? term32(Ms, (float)ldo_s, ldt, eb) : 0.0f;
      const float nr = R > 0 ? term32(R, ldn, (float)l_dr, eb) : 0.0f;
      const float ns = S > 0 ? term32(S, ldn, (float)l_ds, eb) : 0.0f;
      krow.add(((or_ + os_) - nr) - ns);
    }
    if (col_use) {
      int32_t Mr, Ms;
      col_get(Mr, Ms);
      csum_r += Mr;
      csum_s += Ms;
      const float ldt = (float)m.ldout[t_i];
      const float ir = Mr > 0 ? term32(Mr, ldt, (float)ldi_r, eb) : 0.0f;
      const float is = Ms > 0 ? term32(Ms, ldt, (float)ldi_s, eb) : 0.0f;
      const int32_t wr = Mr - c_sum, ws = Ms + c_sum;
      const float nr = wr > 0 ? term32(wr, ldt, (float)l_ir, eb) : 0.0f;
      const float ns = ws > 0 ? term32(ws, ldt, (float)l_is, eb) : 0.0f;
      kcol.add(((ir + is) - nr) - ns);
    }
  }
  sum_r = group_allsum<G>(sum_r);
  sum_s = group_allsum<G>(sum_s);
  csum_r = group_allsum<G>(csum_r);
  csum_s = group_allsum<G>(csum_s);
  const int32_t M_rr = __shfl(corner, 0, G), M_sr = __shfl(corner, 1, G);
  const int32_t M_rs = __shfl(corner, 2, G), M_ss = __shfl(corner, 3, G);
  const int32_t W_r = dout_r - sum_r, W_s = dout_s - sum_s;
  const double Y_r = W_r > 0 ? (ldo_r - l_dr) * (double)W_r : 0.0;
  const double Y_s = W_s > 0 ? (ldo_s - l_ds) * (double)W_s : 0.0;
  const int32_t W2_r = din_r - M_rr - M_sr - csum_r;
  const int32_t W2_s = din_s - M_rs - M_ss - csum_s;
  const double Y2_r = W2_r > 0 ? (ldi_r - l_ir) * (double)W2_r : 0.0;
  const double Y2_s = W2_s > 0 ? (ldi_s - l_is) * (double)W2_s : 0.0;
  double d32 = 0.0, bound = 0.0;
  if constexpr (MIX) {
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) {
      krow.combine(__shfl_down(krow.s, off, G), __shfl_down(krow.c, off, G));
      kcol.combine(__shfl_down(kcol.s, off, G), __shfl_down(kcol.c, off, G));
      eb += __shfl_down(eb, off, G);
    }
    int rej = 0;
    if (q == 0) {
      d32 = nodal_total((double)krow.value(), (double)kcol.value(), Y_r, Y_s, Y2_r, Y2_s);
      rej = mixed_rejects(d32, eb, Y_r, Y_s, Y2_r, Y2_s, bound);
    }
    if (__shfl(rej, 0, G)) {
      if (q == 0) {
        a.dS[i] = 0.0;
        a.err[i] = 0.0f;
        a.near[i] = 0;
      }
      return;
    }
    fp64_pass(drow, dcol, false);
  }
  double lh = 0.0;
  if constexpr (MH) {
    const int32_t a_r = group_allsum<G>((q < no && t_o == r) ? w_o : 0);
    const int32_t a_s = group_allsum<G>((q < no && t_o == s) ? w_o : 0);
    const Hast h{v, r, s, B, M_rs, M_sr, M_ss, M_rr - a_r - c_r + sl, M_rs - a_s + c_r - sl,
                 M_sr + a_r - c_s - sl, tot, dout_r + din_r, dout_s + din_s};
    CAcc<G> fo, fi, bo, bi;
    fo.init();
    fi.init();
    bo.init();
    bi.init();
    if (q < no) {
      double f, b;
      hast_terms(m, h, u_o, t_o, w_o, a_sum, c_of_o, f, b);
      fo.p[0] += f;
      bo.p[0] += b;
    }
    if (q < ni) {
      double f, b;
      hast_terms(m, h, u_i, t_i, w_i, a_of_i, c_sum, f, b);
      fi.p[0] += f;
      bi.p[0] += b;
    }
    const double so = fo.reduce(), si = fi.reduce(), ro = bo.reduce(), ri = bi.reduce();
    if (q == 0) lh = hast_log2(so, si, ro, ri, no + ni);
  }
  if (q == 0) {
    const double d = nodal_total(drow, dcol, Y_r, Y_s, Y2_r, Y2_s);
    if constexpr (MH) {
      const bool acc = mh_accepts(d, lh, a.beta, x.x2);
      a.prop[i] = acc ? s : r;
      a.dS[i] = acc ? d : 0.0;
      return;
    }
    a.dS[i] = d < 0.0 ? d : 0.0;
    if constexpr (MIX) {
      // non-finite delta-S (an empty block's -inf log, the reference's quirk) only ever
      // sends a proposal to the FP64 pass; it is left out of the error statistic
      const double e = __builtin_fabs(d32 - d);
      a.err[i] = e <= 1e30 ? (float)e : 0.0f;
      a.near[i] = d32 > -bound ? 2 : 1;
    }
  }
}

// Proposals of a batch on a unit-weight graph, one lane per vertex (kernels.hpp:877-939,
// 1000-1032): with unit weights the first adjacency entry whose running weight exceeds the
// target is entry `target` itself, so one load finds the neighbour.  Writes prop for every
// vertex and the outcome of those that propose their own block; the fused kernel's teams
// then start from prop (NodalArgs::pre).
__global__ __launch_bounds__(256) void k_nodal_propose(NodalArgs a) {
  __shared__ uint32_t wc[NCLS], wb[NCLS];
  if (threadIdx.x < NCLS) wc[threadIdx.x] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int cls = -1;
  uint32_t lp = 0;
  int64_t v = 0;
  int32_t rec_o0 = 0, rec_i0 = 0, rec_no = 0, rec_ni = 0;
  if (i < a.BS) {
    v = a.lo + i;
    const int32_t B = a.m.B;
    const int32_t r = a.lab[v];
    const int32_t o0 = a.optr[v], no = a.optr[v + 1] - o0;
    const int32_t i0 = a.iptr[v], ni = a.iptr[v + 1] - i0;
    rec_o0 = o0; rec_i0 = i0; rec_no = no; rec_ni = ni;
    const Philox x =
        philox((uint32_t)v, nodal_phase(a), a.key.chain, TAG_NODE, a.key.k0, a.key.k1);
    const uint32_t tot = (uint32_t)(no + ni);
    int32_t s;
    if (tot == 0) {
      s = (int32_t)mulhi(x.x0, B);
    } else {
      const uint32_t target = mulhi(x.x1, tot);
      const int32_t nbr = target < (uint32_t)no ? a.oadj[o0 + target] : a.iadj[i0 + target - no];
      const int32_t u = a.lab[nbr];
      s = a.coin[u] ? (int32_t)mulhi(x.x0, B) : a.sof[u];
    }
    a.prop[i] = s;
    if (s == r) {
      a.dS[i] = 0.0;
      if (a.err) {
        a.err[i] = 0.0f;
        a.near[i] = 0;
      }
    } else {
      // the neighbour-count class of the fused kernel's teams
      const int32_t d = no > ni ? no : ni;
      cls = d <= a.cb16 ? 0 : d <= a.cb32 ? 1 : d <= a.cbwave ? 2 : 3;
      lp = atomicAdd(&wc[cls], 1u);
    }
  }
  if (a.evstat) {
    // workgroup totals in LDS, then one pair of global atomics per workgroup, spread over
    // EV_SLOTS slot pairs
    __shared__ uint32_t ev[2];
    if (threadIdx.x < 2) ev[threadIdx.x] = 0;
    __syncthreads();
    if (cls >= 0) {
      atomicAdd(&ev[0], 1u);
      atomicAdd(&ev[1], (uint32_t)((a.optr[v + 1] - a.optr[v]) + (a.iptr[v + 1] - a.iptr[v])));
    }
    __syncthreads();
    if (threadIdx.x < 2 && ev[threadIdx.x])
      atomicAdd(&a.evstat[2 * (blockIdx.x % EV_SLOTS) + threadIdx.x],
                (unsigned long long)ev[threadIdx.x]);
  }
  __syncthreads();
  if (threadIdx.x < NCLS)
    wb[threadIdx.x] = wc[threadIdx.x] ? atomicAdd((uint32_t *)&a.dcnt[threadIdx.x],
                                                  wc[threadIdx.x]) : 0u;
  __syncthreads();
  if (cls >= 0)
    a.dlist[(int64_t)cls * a.BS + wb[cls] + lp] =
        make_int4((int32_t)v, rec_o0, rec_i0, cls < NCLS - 1 ? (rec_no | rec_ni << 16) : 0);
}

// Vertex lists of one batch by class, and the workgroups of each class in the fused launch
struct NodalLists {
  const int32_t *list[NCLS];  // host-built lists (vertex ids), or
  const int4 *rec[NCLS];      // the proposal kernel's records (NodalArgs::dlist), or null
  int32_t n[NCLS];
  int32_t nb[NCLS];
  uint32_t big_lds;
  const int32_t *dn;  // device counts of the lists (proposal kernel's), or null: n
};

// One launch per batch: the largest vertices first (one workgroup each), then 4 vertices
// per workgroup (a wave each, 6 KB of LDS tables per wave), then groups of 32 and 16 lanes.
template <bool MIX, int VAR>
GSAP_D void nodal_fused(const NodalArgs &a, const NodalLists &L, double *lds_d) {
  // listed vertices of each class (the grid is sized for all of a batch's vertices)
  auto cnt = [&](int k) { return L.dn ? L.dn[k] : L.n[k]; };
  int32_t b = (int32_t)blockIdx.x;
  if (b < L.nb[3]) {
    if (b >= cnt(3)) return;
    log2_tables_to_lds();
    nodal_big<MIX, VAR>(a, L.rec[3] ? (int64_t)L.rec[3][b].x : (int64_t)L.list[3][b], lds_d,
                        L.big_lds);
    return;
  }
  b -= L.nb[3];
  if (b < L.nb[2]) {
    const int32_t n2 = cnt(2);
    if (b * 4 >= n2) return;
    log2_tables_to_lds();
    const int32_t k = b * 4 + (int32_t)(threadIdx.x >> 6);
    if (k < n2)
      nodal_wave<MIX, VAR>(a, L.rec[2] ? vhead_rec(L.rec[2][k]) : vhead_load(a, L.list[2][k]),
                           L.rec[2] != nullptr,
                           reinterpret_cast<int32_t *>(lds_d) + (threadIdx.x >> 6) * 6 * LCAP,
                           (uint32_t)LBITS);
    return;
  }
  b -= L.nb[2];
  if (b < L.nb[1]) {
    const int32_t n1 = cnt(1);
    if (b * 8 >= n1) return;
    log2_tables_to_lds();
    nodal_small<32, MIX, VAR>(a, L.list[1], L.rec[1], n1, b);
    return;
  }
  b -= L.nb[1];
  const int32_t n0 = cnt(0);
  if (b * 16 >= n0) return;
  log2_tables_to_lds();
  nodal_small<16, MIX, VAR>(a, L.list[0], L.rec[0], n0, b);
}

template <bool MIX, int VAR>
__global__ __launch_bounds__(256) void k_nodal_fused(NodalArgs a, NodalLists L) {
  extern __shared__ double lds_d[];
  nodal_fused<MIX, VAR>(a, L, lds_d);
}

// The FP64 greedy forms fit 96 registers without spills: five waves per SIMD instead of four
// (the mixed and Metropolis-Hastings forms would spill there).
template <bool MIX, int VAR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_nodal_fused5(
    NodalArgs a, NodalLists L) {
  extern __shared__ double lds_d[];
  nodal_fused<MIX, VAR>(a, L, lds_d);
}

void launch_nodal(const GraphDev &g, LabView lab, const BMView &m, const int32_t *sof,
                  const uint8_t *coin, RngKey key, PhaseRef phase, int64_t lo, int64_t BS,
                  const int32_t *const lists[NCLS], const int32_t counts[NCLS], uint32_t big_lds,
                  int32_t *prop, double *dS, float *err, uint8_t *near, bool exact, bool mh,
                  double beta, const NodalPre &pre, int stages, hipStream_t st) {
  if (BS <= 0) return;
  NodalArgs a;
  a.lo = lo; a.BS = BS;
  a.optr = g.optr; a.oadj = g.oadj; a.ow = g.ow;
  a.iptr = g.iptr; a.iadj = g.iadj; a.iw = g.iw;
  a.kout = g.kout; a.kin = g.kin; a.lab = lab;
  a.m = m; a.sof = sof; a.coin = coin; a.key = key; a.phase = phase;
  a.prop = prop; a.dS = dS;
  a.hoff = g.hoff; a.hbits_o = g.hbits_o; a.hbits_i = g.hbits_i; a.hpool = g.hpool;
  a.err = err; a.near = near;
  a.sloop = g.sloop; a.beta = beta;
  // unit-weight graphs: every proposal first, one lane per vertex, and the vertices that
  // propose another block listed by class for the fused kernel
  a.pre = (g.wmax == 1 && pre.dlist) ? 1 : 0;
  a.dlist = pre.dlist;
  a.dcnt = pre.dcnt;
  a.cb16 = pre.bounds.max16;
  a.cb32 = pre.bounds.max32;
  a.cbwave = pre.bounds.maxwave;
  a.evstat = pre.evstat;
  if (a.pre && (stages & NODAL_PROPOSE))
    k_nodal_propose<<<(unsigned)((BS + 255) / 256), 256, 0, st>>>(a);
  if (!(stages & NODAL_MOVES)) return;
  NodalLists L;
  const int per_block[NCLS] = {16, 8, 4, 1};
  int64_t total = 0;
  for (int k = 0; k < NCLS; ++k) {
    L.list[k] = lists[k];
    L.rec[k] = a.pre ? pre.dlist + (int64_t)k * BS : nullptr;
    L.n[k] = counts[k];
    L.nb[k] = (counts[k] + per_block[k] - 1) / per_block[k];
    total += L.nb[k];
  }
  L.big_lds = big_lds;
  L.dn = a.pre ? pre.dcnt : nullptr;
  if (total == 0) return;
  size_t lds = (size_t)4 * 6 * 4 * LCAP;  // four waves' tables
  if (L.nb[3] > 0 && big_lds > lds) lds = big_lds;
  const dim3 grid((unsigned)total);
  // the mixed form supports the greedy rule only (the host rejects mixed with MH)
  if (err && !exact)
    k_nodal_fused<true, 0><<<grid, 256, lds, st>>>(a, L);
  else if (err)
    k_nodal_fused<true, VAR_EXACT><<<grid, 256, lds, st>>>(a, L);
  else if (!exact && !mh)
    k_nodal_fused5<false, 0><<<grid, 256, lds, st>>>(a, L);
  else if (!mh)
    k_nodal_fused5<false, VAR_EXACT><<<grid, 256, lds, st>>>(a, L);
  else if (!exact)
    k_nodal_fused<false, VAR_MH><<<grid, 256, lds, st>>>(a, L);
  else
    k_nodal_fused<false, VAR_EXACT | VAR_MH><<<grid, 256, lds, st>>>(a, L);
}

// Mixed-form statistics of one batch: max |dS32 - dS64| (float bits, order-preserving for
// values >= 0), the count of near-threshold decisions (class 2) and of FP64 evaluations
// (classes 1 and 2).
__global__ __launch_bounds__(256) void k_mix_stats(const float *err, const uint8_t *cls,
                                                   int64_t n, uint32_t *max_bits,
                                                   unsigned long long *count) {
  uint32_t mx = 0;
  uint32_t c_near = 0, c_eval = 0;
  for (int64_t k = blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) {
    const uint32_t b = __builtin_bit_cast(uint32_t, err[k]);
    mx = b > mx ? b : mx;
    c_near += cls[k] == 2;
    c_eval += cls[k] != 0;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const uint32_t y = __shfl_down(mx, off, 64);
    mx = y > mx ? y : mx;
    c_near += __shfl_down(c_near, off, 64);
    c_eval += __shfl_down(c_eval, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (mx) atomicMax(max_bits, mx);
    if (c_near) atomicAdd(count, (unsigned long long)c_near);
    if (c_eval) atomicAdd(count + 1, (unsigned long long)c_eval);
  }
}

void launch_mix_stats(const float *err, const uint8_t *near, int64_t n, uint32_t *max_bits,
                      unsigned long long *count, hipStream_t st) {
  if (n <= 0) return;
  unsigned grid = (unsigned)((n + 255) / 256);
  if (grid > 1024) grid = 1024;
  k_mix_stats<<<grid, 256, 0, st>>>(err, near, n, max_bits, count);
}


// ---------------------------------------------------------- canonical sums
// the sum of the tile sums (ctr, when given, advanced by add afterwards: the sweep's batch
// counter, stream-ordered after the sweep's last read of it)
__global__ __launch_bounds__(64) void k_final_sum(const double *tiles, const int32_t *n_dev,
                                                  int64_t n_host, double *out, uint32_t *ctr,
                                                  uint32_t add) {
  const int64_t n = n_dev ? (int64_t)*n_dev : n_host;
  const int64_t nt = (n + 1023) / 1024;
  const int lane = threadIdx.x;
  double p = 0.0;
  for (int64_t k = lane; k < nt; k += 64) p += tiles[k];
  p = wave_sum(p);
  if (lane == 0) {
    *out = p;
    if (ctr) *ctr += add;
  }
}

// Labels of a nodal batch and the tile sums of its delta-S, one 1024-entry tile per
// 256-thread workgroup (four entries per thread, then wave 0 sums the tile from LDS in the
// canonical order): greedy (all = 0) takes prop where dS < 0, Metropolis-Hastings (all = 1)
// takes prop; olab keeps the previous labels, lab16 (when given) follows the changed ones.
// DEG (a batch whose CSR rebuild is deferred): a moved vertex's weighted out- and in-degree
// leave its old block and enter its new one, summed per workgroup in LDS (DEG = 1, 2B words,
// one global atomic per touched block) or added with global atomics (DEG = 2, large B).
template <int DEG>
__global__ __launch_bounds__(256) void k_apply_tiles(int32_t *lab, const int32_t *prop,
                                                     const double *dS, int64_t lo, int64_t BS,
                                                     int all, int32_t *olab, uint16_t *lab16,
                                                     double *tiles, DegMoves dm) {
  __shared__ double xs[1024];
  extern __shared__ int32_t sdeg[];
  if constexpr (DEG == 1) {
    for (int32_t k = threadIdx.x; k < 2 * dm.B; k += 256) sdeg[k] = 0;
    __syncthreads();
  }
  const int64_t t0 = (int64_t)blockIdx.x * 1024;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = (int)threadIdx.x + 256 * j;
    const int64_t i = t0 + k;
    double d = 0.0;
    if (i < BS) {
      d = dS[i];
      const int32_t old = lab[lo + i];
      if (olab) olab[i] = old;
      if (all || d < 0.0) {
        const int32_t nl = prop[i];
        lab[lo + i] = nl;
        if (lab16) lab16[lo + i] = (uint16_t)nl;
        if (DEG && nl != old) {
          const int64_t x = lo + i;
          const int32_t ko = dm.kout ? dm.kout[x] : dm.optr[x + 1] - dm.optr[x];
          const int32_t ki = dm.kin ? dm.kin[x] : dm.iptr[x + 1] - dm.iptr[x];
          int32_t *po = DEG == 1 ? sdeg : dm.dout, *pi = DEG == 1 ? sdeg + dm.B : dm.din;
          if (ko) {
            atomicSub(&po[old], ko);
            atomicAdd(&po[nl], ko);
          }
          if (ki) {
            atomicSub(&pi[old], ki);
            atomicAdd(&pi[nl], ki);
          }
        }
      }
    }
    xs[k] = d;
  }
  __syncthreads();
  if constexpr (DEG == 1) {
    for (int32_t b = threadIdx.x; b < dm.B; b += 256) {
      if (sdeg[b]) atomicAdd(&dm.dout[b], sdeg[b]);
      if (sdeg[dm.B + b]) atomicAdd(&dm.din[b], sdeg[dm.B + b]);
    }
  }
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  double p = 0.0;
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (t0 + lane + 64 * j < BS) p += xs[lane + 64 * j];
  p = wave_sum(p);
  if (lane == 0) tiles[blockIdx.x] = p;
}

constexpr int DEG_LDS_B = 4096;  // largest B whose degree moves are summed in LDS

void launch_apply_sum(int32_t *lab, const int32_t *prop, const double *dS, int64_t lo, int64_t BS,
                      bool all, int32_t *olab, uint16_t *lab16, double *tiles, double *out,
                      const DegMoves &dm, hipStream_t st) {
  const int64_t nt = (BS + 1023) / 1024;
  const int a = all ? 1 : 0;
  if (nt > 0) {
    if (!dm.dout)
      k_apply_tiles<0><<<(unsigned)nt, 256, 0, st>>>(lab, prop, dS, lo, BS, a, olab, lab16, tiles,
                                                     dm);
    else if (dm.B <= DEG_LDS_B)
      k_apply_tiles<1><<<(unsigned)nt, 256, (size_t)8 * dm.B, st>>>(lab, prop, dS, lo, BS, a,
                                                                     olab, lab16, tiles, dm);
    else
      k_apply_tiles<2><<<(unsigned)nt, 256, 0, st>>>(lab, prop, dS, lo, BS, a, olab, lab16, tiles,
                                                     dm);
  }
  if (out) k_final_sum<<<1, 64, 0, st>>>(tiles, nullptr, BS, out, nullptr, 0);
}


// data term of every block-matrix entry in CSR order (kernels.hpp:1066-1093), summed per
// 1024-entry tile (one wave each)
__global__ __launch_bounds__(64) void k_mdl_tiles(BMView m, const int32_t *erow, double *tiles) {
  const int64_t n = m.rowptr[m.B];
  const int64_t t0 = (int64_t)blockIdx.x * 1024;
  if (t0 >= n) return;
  log2_tables_to_lds();
  const int lane = threadIdx.x;
  double p = 0.0;
#pragma unroll 4
  for (int j = 0; j < 16; ++j) {
    const int64_t k = t0 + lane + 64 * j;
    if (k < n) p += term_s(m.val[k], m.ldout[erow[k]], m.ldin[m.col[k]]);
  }
  p = wave_sum(p);
  if (lane == 0) tiles[blockIdx.x] = p;
}

void launch_mdl_sum(const BMView &m, const int32_t *erow, int64_t n_max, double *tiles,
                    double *out, uint32_t *ctr, uint32_t add, hipStream_t st) {
  const int64_t nt = (n_max + 1023) / 1024;
  if (nt > 0) k_mdl_tiles<<<(unsigned)nt, 64, 0, st>>>(m, erow, tiles);
  k_final_sum<<<1, 64, 0, st>>>(tiles, m.rowptr + m.B, 0, out, ctr, add);
}

// The sums that end a sweep, one wave each: workgroup 0 the description-length data term
// (its tiles at tiles[0..), n from nnz_dev; it also advances ctr by add), workgroup 1 + b
// batch b's delta-S (its tiles at tiles + boff + b * bstride, n = BS).  out[0], out[1 + b].
__global__ __launch_bounds__(64) void k_sweep_final(const double *tiles, const int32_t *nnz_dev,
                                                    int64_t boff, int64_t bstride, int64_t BS,
                                                    double *out, uint32_t *ctr, uint32_t add) {
  const int x = blockIdx.x;
  const int64_t n = x == 0 ? (int64_t)*nnz_dev : BS;
  const double *t = x == 0 ? tiles : tiles + boff + (int64_t)(x - 1) * bstride;
  const int64_t nt = (n + 1023) / 1024;
  const int lane = threadIdx.x;
  double p = 0.0;
  for (int64_t k = lane; k < nt; k += 64) p += t[k];
  p = wave_sum(p);
  if (lane == 0) {
    out[x] = p;
    if (x == 0 && ctr) *ctr += add;
  }
}

void launch_sweep_sums(const BMView &m, const int32_t *erow, int64_t n_max, double *tiles,
                       int64_t boff, int64_t bstride, int32_t nb, int64_t BS, double *out,
                       uint32_t *ctr, uint32_t add, hipStream_t st) {
  const int64_t nt = (n_max + 1023) / 1024;
  if (nt > 0) k_mdl_tiles<<<(unsigned)nt, 64, 0, st>>>(m, erow, tiles);
  k_sweep_final<<<(unsigned)(nb + 1), 64, 0, st>>>(tiles, m.rowptr + m.B, boff, bstride, BS, out,
                                                   ctr, add);
}

// One thread: n doubles and the entry count into host memory, a system-scope fence, then
// the sequence number, which the host polls.
__global__ __launch_bounds__(64) void k_publish(const double *res, int32_t n, const int32_t *nnz,
                                                double *hv, int32_t *hnnz, uint32_t *hseq,
                                                uint32_t seq) {
  if (threadIdx.x != 0) return;
  for (int32_t i = 0; i < n; ++i) hv[i] = res[i];
  *hnnz = *nnz;
  __threadfence_system();
  __hip_atomic_store(hseq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void launch_publish(const double *res, int32_t n, const int32_t *nnz, double *hv, int32_t *hnnz,
                    uint32_t *hseq, uint32_t seq, hipStream_t st) {
  k_publish<<<1, 64, 0, st>>>(res, n, nnz, hv, hnnz, hseq, seq);
}


// ------------------------------------------------------------- block merges
// One group of G lanes per proposal (r, i): the proposal (kernels.hpp:726-856, every lane
// computes it) and the merge delta-S (kernels.hpp:143-449, s-side column fixed).  Exact
// form: the merged row and column hold t outside {r, s}, the four entries of rows and
// columns r, s fold into one with the merged degrees, and a proposal of r itself gets the
// largest double.
template <int G>
GSAP_D void merge_one(const BMView &m, const double *OUT, RngKey key, uint32_t phase,
                      int32_t np, int64_t idx, int32_t *prop, double *dSout, int exact) {
  const int32_t B = m.B;
  const int q = threadIdx.x & (G - 1);
  const int32_t r = (int32_t)(idx / np);
  const Philox x = philox((uint32_t)idx, phase, key.chain, TAG_MERGE, key.k0, key.k1);
  int32_t s;
  const int64_t deg = (int64_t)m.dout[r] + m.din[r];
  if (deg == 0) {
    s = (int32_t)mulhi(x.x0, B);
  } else {
    const int32_t u = group_walk<G>(m, r, mulhi(x.x1, deg), -1, 0);
    const int64_t degu = (int64_t)m.dout[u] + m.din[u];
    if (coin_random(x.x3, degu, B)) s = (int32_t)mulhi(x.x0, B);
    else s = group_walk<G>(m, u, mulhi(x.x2, degu), r, 0);
  }
  if (exact && s == r) {
    if (q == 0) {
      prop[idx] = s;
      dSout[idx] = 1.7976931348623157e308;
    }
    return;
  }
  // Column r: its old terms and the merged column's terms over its entries (t not in
  // {r, s}); column s: its old terms and the merged column's terms of entries absent from
  // column r.  Then the merged row: row r's entries (M_rt + M_st), and row s's entries
  // absent from row r.  Four entries per lane (two in groups of 8) are loaded before any is
  // summed.
  constexpr int P = G >= 16 ? 4 : 2;
  const double ldo = clog2_s((double)((int64_t)m.dout[r] + m.dout[s]));
  const double ldi = clog2_s((double)((int64_t)m.din[r] + m.din[s]));
  double inx_r, inx_s, ia_, ib_;
  {
    CAcc<G> ox, na;
    ox.init();
    na.init();
    const int32_t b0 = m.colptr[r], n = m.colptr[r + 1] - b0;
    const double li = m.ldin[r];
    seg_for3<G, P>(
        n, [&](int32_t k) { return make_int2(m.row[b0 + k], m.cval[b0 + k]); },
        [&](int2 e) { return GL{bm_col_get(m, s, e.x), m.ldout[e.x]}; },
        [&](int2 e, GL g, int j) {
          if (e.x == r || e.x == s) return;
          ox.p[j] += term_s(e.y, g.l, li);
          na.p[j] += term_s((int64_t)e.y + g.g, g.l, ldi);
        });
    inx_r = ox.reduce();
    ia_ = na.reduce();
  }
  {
    CAcc<G> ox, nb;
    ox.init();
    nb.init();
    const int32_t b0 = m.colptr[s], n = m.colptr[s + 1] - b0;
    const double li = m.ldin[s];
    seg_for3<G, P>(
        n, [&](int32_t k) { return make_int2(m.row[b0 + k], m.cval[b0 + k]); },
        [&](int2 e) { return GL{s != r ? bm_col_get(m, r, e.x) : 1, m.ldout[e.x]}; },
        [&](int2 e, GL g, int j) {
          if (e.x == r || e.x == s) return;
          ox.p[j] += term_s(e.y, g.l, li);
          if (g.g == 0) nb.p[j] += term_s(e.y, g.l, ldi);
        });
    inx_s = ox.reduce();
    ib_ = nb.reduce();
  }
  double new_out;
  {
    CAcc<G> na, nb;
    na.init();
    nb.init();
    const int32_t b0 = m.rowptr[r], n = m.rowptr[r + 1] - b0;
    seg_for3<G, P>(
        n, [&](int32_t k) { return make_int2(m.col[b0 + k], m.val[b0 + k]); },
        [&](int2 e) { return GL{bm_row_get(m, s, e.x), m.ldin[e.x]}; },
        [&](int2 e, GL g, int j) {
          if (exact && (e.x == r || e.x == s)) return;
          na.p[j] += term_s((int64_t)e.y + g.g, g.l, ldo);
        });
    if (s != r) {
      const int32_t c0 = m.rowptr[s], c = m.rowptr[s + 1] - c0;
      seg_for3<G, P>(
          c, [&](int32_t k) { return make_int2(m.col[c0 + k], m.val[c0 + k]); },
          [&](int2 e) { return GL{bm_row_get(m, r, e.x), m.ldin[e.x]}; },
          [&](int2 e, GL g, int j) {
            if (exact && (e.x == r || e.x == s)) return;
            if (g.g == 0) nb.p[j] += term_s(e.y, g.l, ldo);
          });
    }
    const double a_ = na.reduce();
    const double b_ = nb.reduce();
    new_out = a_ + b_;
  }
  if (exact) {
    // the folded corner M_rr + M_rs + M_sr + M_ss
    const int64_t X = (int64_t)bm_row_get(m, r, r) + bm_row_get(m, r, s) + bm_row_get(m, s, r) +
                      bm_row_get(m, s, s);
    new_out = new_out + (X > 0 ? term_s(X, ldi, ldo) : 0.0);
  }
  const double new_in = ia_ + ib_;
  if (q == 0) {
    double d = OUT[r] + OUT[s];
    d = d + (inx_r + inx_s);
    d = d - new_out;
    d = d - new_in;
    prop[idx] = s;
    dSout[idx] = d;
  }
}

template <int G>
__global__ __launch_bounds__(256) void k_merge(BMView m, const double *OUT, RngKey key,
                                               uint32_t phase, int32_t np, int32_t *prop,
                                               double *dSout, int exact) {
  log2_tables_to_lds();
  const int64_t idx = (int64_t)blockIdx.x * (256 / G) + threadIdx.x / G;
  if (idx >= (int64_t)m.B * np) return;
  merge_one<G>(m, OUT, key, phase, np, idx, prop, dSout, exact);
}

// blocks whose row and column hold more than cap entries, in any order
__global__ void k_big_blocks(BMView m, int32_t cap, int32_t *list, uint32_t *count) {
  const int32_t r = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= m.B) return;
  if ((m.rowptr[r + 1] - m.rowptr[r]) + (m.colptr[r + 1] - m.colptr[r]) > cap)
    list[atomicAdd(count, 1u)] = r;
}

// one wave per proposal of the listed blocks (*count of them, written by k_big_blocks on the
// same stream), waves striding over the proposals
__global__ __launch_bounds__(256) void k_merge_list(BMView m, const double *OUT, RngKey key,
                                                    uint32_t phase, int32_t np, int32_t *prop,
                                                    double *dSout, int exact,
                                                    const int32_t *list, const uint32_t *count) {
  const int64_t n = (int64_t)*count * np;
  if ((int64_t)blockIdx.x * 4 >= n) return;
  log2_tables_to_lds();
  for (int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); j < n;
       j += (int64_t)gridDim.x * 4) {
    const int32_t r = list[j / np];
    merge_one<64>(m, OUT, key, phase, np, (int64_t)r * np + j % np, prop, dSout, exact);
  }
}

// position of t in the sorted keys[0, n) (LDS), or -1
GSAP_D int32_t lds_find(const int32_t *keys, int32_t n, int32_t t) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (keys[mid] < t) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && keys[lo] == t) ? lo : -1;
}

// Block merges with row r and column r staged in LDS: one 256-thread workgroup per block r,
// one wave per proposal (r, i), i = wave, wave + 4, ...  Row s is streamed once: each of its
// entries is found in row r by binary search in LDS (absent: a row-s-only term of the merged
// row; present: its value M_st lands at that position of the wave's match array, read by the
// row-r terms); column s likewise against column r.  The random lookups of merge_one become
// LDS work.  Same positions, terms and canonical sums as merge_one, hence the same bits; the
// proposals of a block whose row and column exceed CAP entries are left to k_merge_list.
// (k_merge_blk_g: the same with groups of G lanes, for short rows.)
template <int CAP>
__global__ __launch_bounds__(256) void k_merge_blk_w(BMView m, const double *OUT, RngKey key,
                                                   uint32_t phase, int32_t np, int32_t *prop,
                                                   double *dSout, int exact) {
  __shared__ int32_t sk[CAP], sv[CAP];  // row r (columns, values), then column r (rows, values)
  __shared__ int32_t mt[4][CAP];        // per wave: M_st at row r's positions, M_ts at column r's
  log2_tables_to_lds();
  const int32_t r = (int32_t)blockIdx.x;
  const int32_t B = m.B;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int32_t r0 = m.rowptr[r], nr = m.rowptr[r + 1] - r0;
  const int32_t c0 = m.colptr[r], nc = m.colptr[r + 1] - c0;
  const int32_t n = nr + nc;
  if (n > CAP) return;  // listed by k_big_blocks, one wave per proposal
  for (int32_t k = threadIdx.x; k < nr; k += 256) {
    sk[k] = m.col[r0 + k];
    sv[k] = m.val[r0 + k];
  }
  for (int32_t k = threadIdx.x; k < nc; k += 256) {
    sk[nr + k] = m.row[c0 + k];
    sv[nr + k] = m.cval[c0 + k];
  }
  __syncthreads();
  const int32_t *rk = sk, *ck = sk + nr;
  int32_t *mw = mt[wv];
  const int64_t deg = (int64_t)m.dout[r] + m.din[r];
  const double li_r = m.ldin[r];
  for (int32_t i = wv; i < np; i += 4) {
    const int64_t idx = (int64_t)r * np + i;
    const Philox x = philox((uint32_t)idx, phase, key.chain, TAG_MERGE, key.k0, key.k1);
    int32_t s;
    if (deg == 0) {
      s = (int32_t)mulhi(x.x0, B);
    } else {
      // u: first entry of row r then column r (staged) whose running weight exceeds target
      const uint64_t target = mulhi(x.x1, deg);
      int32_t u = 0;
      uint64_t base = 0;
      for (int32_t c = 0; c < n; c += 64) {
        const int32_t k = c + lane;
        uint64_t incl = k < n ? (uint64_t)sv[k] : 0;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const uint64_t y = __shfl_up(incl, off, 64);
          if (lane >= off) incl += y;
        }
        const uint64_t bal = __ballot(k < n && base + incl > target);
        if (bal) {
          u = sk[c + __ffsll((long long)bal) - 1];
          break;
        }
        base += __shfl(incl, 63, 64);
      }
      const int64_t degu = (int64_t)m.dout[u] + m.din[u];
      if (coin_random(x.x3, degu, B)) s = (int32_t)mulhi(x.x0, B);
      else s = group_walk<64>(m, u, mulhi(x.x2, degu), r, 0);
    }
    if (exact && s == r) {
      if (lane == 0) {
        prop[idx] = s;
        dSout[idx] = 1.7976931348623157e308;
      }
      continue;
    }
    wave_lds_sync();  // the previous proposal's reads of the match array are done
    for (int32_t k = lane; k < n; k += 64) mw[k] = s == r ? sv[k] : 0;
    wave_lds_sync();
    const double ldo = clog2_s((double)((int64_t)m.dout[r] + m.dout[s]));
    const double ldi = clog2_s((double)((int64_t)m.din[r] + m.din[s]));
    const double li_s = m.ldin[s];
    // row s: the terms of entries absent from row r, and M_st at row r's positions
    double p_nb = 0.0;
    if (s != r) {
      const int32_t s0 = m.rowptr[s], ns = m.rowptr[s + 1] - s0;
      for (int32_t k = lane; k < ns; k += 64) {
        const int32_t t = m.col[s0 + k], v = m.val[s0 + k];
        const int32_t pos = lds_find(rk, nr, t);
        if (pos >= 0) mw[pos] = v;
        else if (!(exact && (t == r || t == s))) p_nb += term_s(v, m.ldin[t], ldo);
      }
    }
    // column s: its old terms (t not in {r, s}), the terms of entries absent from column r,
    // and M_ts at column r's positions
    double p_inxs = 0.0, p_ib = 0.0;
    {
      const int32_t cs0 = m.colptr[s], ncs = m.colptr[s + 1] - cs0;
      for (int32_t k = lane; k < ncs; k += 64) {
        const int32_t t = m.row[cs0 + k], v = m.cval[cs0 + k];
        const bool ex = t == r || t == s;
        if (!ex) p_inxs += term_s(v, m.ldout[t], li_s);
        if (s != r) {
          const int32_t pos = lds_find(ck, nc, t);
          if (pos >= 0) mw[nr + pos] = v;
          else if (!ex) p_ib += term_s(v, m.ldout[t], ldi);
        }
      }
    }
    wave_lds_sync();
    // merged row over row r's entries, old column r and merged column over column r's
    double p_na = 0.0, p_inxr = 0.0, p_ia = 0.0;
    for (int32_t k = lane; k < nr; k += 64) {
      const int32_t t = rk[k];
      if (exact && (t == r || t == s)) continue;
      p_na += term_s((int64_t)sv[k] + mw[k], m.ldin[t], ldo);
    }
    for (int32_t k = lane; k < nc; k += 64) {
      const int32_t t = ck[k];
      if (t == r || t == s) continue;
      const int32_t v = sv[nr + k];
      p_inxr += term_s(v, m.ldout[t], li_r);
      p_ia += term_s((int64_t)v + mw[nr + k], m.ldout[t], ldi);
    }
    const double inx_r = wave_sum(p_inxr), inx_s = wave_sum(p_inxs);
    const double a_ = wave_sum(p_na), b_ = wave_sum(p_nb);
    const double ia_ = wave_sum(p_ia), ib_ = wave_sum(p_ib);
    if (lane == 0) {
      double new_out = a_ + b_;
      if (exact) {
        const int64_t X = (int64_t)bm_row_get(m, r, r) + bm_row_get(m, r, s) +
                          bm_row_get(m, s, r) + bm_row_get(m, s, s);
        new_out = new_out + (X > 0 ? term_s(X, ldi, ldo) : 0.0);
      }
      const double new_in = ia_ + ib_;
      double d = OUT[r] + OUT[s];
      d = d + (inx_r + inx_s);
      d = d - new_out;
      d = d - new_in;
      prop[idx] = s;
      dSout[idx] = d;
    }
  }
}

// k_merge_blk_w with one group of G lanes per proposal (r, i), i = group, group + NT / G,
// ...: groups of 16 lanes keep several proposals of a block with short rows in flight in one
// wave.  Same positions, terms and canonical sums, hence the same bits.
template <int G, int CAP, int NT, int P, int W>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(W))) void k_merge_blk_g(BMView m, const double *OUT, RngKey key,
                                                   uint32_t phase, int32_t np, int32_t *prop,
                                                   double *dSout, int exact) {
  constexpr int NG = NT / G;
  __shared__ int32_t sk[CAP], sv[CAP];  // row r (columns, values), then column r (rows, values)
  __shared__ int32_t mt[NG][CAP];       // per group: M_st at row r's positions, M_ts at column r's
  log2_tables_to_lds();
  const int32_t r = (int32_t)blockIdx.x;
  const int32_t B = m.B;
  const int grp = threadIdx.x / G, q = threadIdx.x & (G - 1);
  const int gbase = (threadIdx.x & 63) & ~(G - 1);
  const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
  const int32_t r0 = m.rowptr[r], nr = m.rowptr[r + 1] - r0;
  const int32_t c0 = m.colptr[r], nc = m.colptr[r + 1] - c0;
  const int32_t n = nr + nc;
  if (n > CAP) return;  // listed by k_big_blocks, one wave per proposal
  for (int32_t k = threadIdx.x; k < nr; k += NT) {
    sk[k] = m.col[r0 + k];
    sv[k] = m.val[r0 + k];
  }
  for (int32_t k = threadIdx.x; k < nc; k += NT) {
    sk[nr + k] = m.row[c0 + k];
    sv[nr + k] = m.cval[c0 + k];
  }
  __syncthreads();
  const int32_t *rk = sk, *ck = sk + nr;
  int32_t *mw = mt[grp];
  const int64_t deg = (int64_t)m.dout[r] + m.din[r];
  const double li_r = m.ldin[r];
  for (int32_t i = grp; i < np; i += NG) {
    const int64_t idx = (int64_t)r * np + i;
    const Philox x = philox((uint32_t)idx, phase, key.chain, TAG_MERGE, key.k0, key.k1);
    int32_t s;
    if (deg == 0) {
      s = (int32_t)mulhi(x.x0, B);
    } else {
      // u: first entry of row r then column r (staged) whose running weight exceeds target
      const uint64_t target = mulhi(x.x1, deg);
      int32_t u = 0;
      uint64_t base = 0;
      for (int32_t c = 0; c < n; c += G) {
        const int32_t k = c + q;
        uint64_t incl = k < n ? (uint64_t)sv[k] : 0;
#pragma unroll
        for (int off = 1; off < G; off <<= 1) {
          const uint64_t y = __shfl_up(incl, off, G);
          if (q >= off) incl += y;
        }
        const uint64_t bal = (__ballot(k < n && base + incl > target) >> gbase) & gmask;
        if (bal) {
          u = sk[c + __ffsll((long long)bal) - 1];
          break;
        }
        base += __shfl(incl, G - 1, G);
      }
      const int64_t degu = (int64_t)m.dout[u] + m.din[u];
      if (coin_random(x.x3, degu, B)) s = (int32_t)mulhi(x.x0, B);
      else s = group_walk<G>(m, u, mulhi(x.x2, degu), r, 0);
    }
    if (exact && s == r) {
      if (q == 0) {
        prop[idx] = s;
        dSout[idx] = 1.7976931348623157e308;
      }
      continue;
    }
    wave_lds_sync();  // the previous proposal's reads of the match array are done
    for (int32_t k = q; k < n; k += G) mw[k] = s == r ? sv[k] : 0;
    wave_lds_sync();
    const double ldo = clog2_s((double)((int64_t)m.dout[r] + m.dout[s]));
    const double ldi = clog2_s((double)((int64_t)m.din[r] + m.din[s]));
    const double li_s = m.ldin[s];
    // row s: the terms of entries absent from row r, and M_st at row r's positions
    double b_ = 0.0;
    if (s != r) {
      CAcc<G> nb;
      nb.init();
      const int32_t s0 = m.rowptr[s], ns = m.rowptr[s + 1] - s0;
      seg_for3<G, P>(
          ns, [&](int32_t k) { return make_int2(m.col[s0 + k], m.val[s0 + k]); },
          [&](int2 e) { return m.ldin[e.x]; },
          [&](int2 e, double l, int j) {
            const int32_t pos = lds_find(rk, nr, e.x);
            if (pos >= 0) mw[pos] = e.y;
            else if (!(exact && (e.x == r || e.x == s))) nb.p[j] += term_s(e.y, l, ldo);
          });
      b_ = nb.reduce();
    }
    // column s: its old terms (t not in {r, s}), the terms of entries absent from column r,
    // and M_ts at column r's positions
    double inx_s, ib_;
    {
      CAcc<G> ox, nb;
      ox.init();
      nb.init();
      const int32_t cs0 = m.colptr[s], ncs = m.colptr[s + 1] - cs0;
      seg_for3<G, P>(
          ncs, [&](int32_t k) { return make_int2(m.row[cs0 + k], m.cval[cs0 + k]); },
          [&](int2 e) { return m.ldout[e.x]; },
          [&](int2 e, double l, int j) {
            const bool ex = e.x == r || e.x == s;
            if (!ex) ox.p[j] += term_s(e.y, l, li_s);
            if (s != r) {
              const int32_t pos = lds_find(ck, nc, e.x);
              if (pos >= 0) mw[nr + pos] = e.y;
              else if (!ex) nb.p[j] += term_s(e.y, l, ldi);
            }
          });
      inx_s = ox.reduce();
      ib_ = nb.reduce();
    }
    wave_lds_sync();
    // merged row over row r's entries, old column r and merged column over column r's
    double a_;
    {
      CAcc<G> na;
      na.init();
      seg_for3<G, P>(
          nr, [&](int32_t k) { return make_int2(rk[k], k); },
          [&](int2 e) { return m.ldin[e.x]; },
          [&](int2 e, double l, int j) {
            if (exact && (e.x == r || e.x == s)) return;
            na.p[j] += term_s((int64_t)sv[e.y] + mw[e.y], l, ldo);
          });
      a_ = na.reduce();
    }
    double inx_r, ia_;
    {
      CAcc<G> ox, na;
      ox.init();
      na.init();
      seg_for3<G, P>(
          nc, [&](int32_t k) { return make_int2(ck[k], k); },
          [&](int2 e) { return m.ldout[e.x]; },
          [&](int2 e, double l, int j) {
            if (e.x == r || e.x == s) return;
            const int32_t v = sv[nr + e.y];
            ox.p[j] += term_s(v, l, li_r);
            na.p[j] += term_s((int64_t)v + mw[nr + e.y], l, ldi);
          });
      inx_r = ox.reduce();
      ia_ = na.reduce();
    }
    if (q == 0) {
      double new_out = a_ + b_;
      if (exact) {
        const int64_t X = (int64_t)bm_row_get(m, r, r) + bm_row_get(m, r, s) +
                          bm_row_get(m, s, r) + bm_row_get(m, s, s);
        new_out = new_out + (X > 0 ? term_s(X, ldi, ldo) : 0.0);
      }
      const double new_in = ia_ + ib_;
      double d = OUT[r] + OUT[s];
      d = d + (inx_r + inx_s);
      d = d - new_out;
      d = d - new_in;
      prop[idx] = s;
      dSout[idx] = d;
    }
  }
}

void launch_merge(const BMView &m, const double *OUT, RngKey key, uint32_t phase, int32_t np,
                  int32_t *prop, double *dS, int64_t nnz, bool exact, int32_t *big_list,
                  uint32_t *big_count, hipStream_t st) {
  const int64_t n = (int64_t)m.B * np;
  const int32_t avg_row = (int32_t)(m.B > 0 ? (nnz + m.B - 1) / m.B : 0);
  const int ex = exact ? 1 : 0;
  if (avg_row > 12 && m.B >= 4096 && big_list) {
    // long rows and enough blocks to fill the device: a workgroup per block with its row and
    // column in LDS (short rows: groups of 16 lanes, twelve proposals in flight per
    // workgroup, 80 registers at six waves per SIMD; longer rows: a wave per proposal); the
    // few blocks beyond the staging arrays get a wave per proposal, in a grid sized for the
    // most such blocks the entry count allows (row + column > cap entries each), whose
    // waves read the listed count on the device
    int32_t cap;
    if (avg_row <= 32) {
      cap = 256;
      k_merge_blk_g<16, 256, 192, 1, 6><<<(unsigned)m.B, 192, 0, st>>>(m, OUT, key, phase, np,
                                                                        prop, dS, ex);
    } else if (avg_row <= 40) {
      cap = 384;
      k_merge_blk_g<16, 384, 192, 1, 6><<<(unsigned)m.B, 192, 0, st>>>(m, OUT, key, phase, np,
                                                                        prop, dS, ex);
    } else if (avg_row <= 96) {
      cap = 512;
      k_merge_blk_w<512><<<(unsigned)m.B, 256, 0, st>>>(m, OUT, key, phase, np, prop, dS, ex);
    } else {
      cap = 1024;
      k_merge_blk_w<1024><<<(unsigned)m.B, 256, 0, st>>>(m, OUT, key, phase, np, prop, dS, ex);
    }
    if (hipMemsetAsync(big_count, 0, 4, st) != hipSuccess) return;
    k_big_blocks<<<grid_for(m.B, 256), 256, 0, st>>>(m, cap, big_list, big_count);
    // (the waves stride over all listed proposals, so any grid size is correct)
    const int64_t most = std::min<int64_t>(m.B, 2 * nnz / (cap + 1));
    k_merge_list<<<grid_for(std::min<int64_t>(most * np, 65536), 4), 256, 0, st>>>(
        m, OUT, key, phase, np, prop, dS, ex, big_list, big_count);
    return;
  }
  if (avg_row <= 12)
    k_merge<8><<<grid_for(n, 32), 256, 0, st>>>(m, OUT, key, phase, np, prop, dS, ex);
  else if (avg_row <= 24)
    k_merge<16><<<grid_for(n, 16), 256, 0, st>>>(m, OUT, key, phase, np, prop, dS, ex);
  else if (avg_row <= 48)
    k_merge<32><<<grid_for(n, 8), 256, 0, st>>>(m, OUT, key, phase, np, prop, dS, ex);
  else
    k_merge<64><<<grid_for(n, 4), 256, 0, st>>>(m, OUT, key, phase, np, prop, dS, ex);
}

// find_best_warp (kernels.hpp:452-491): 32 lanes padded with (+MAX, 0), shuffle-down
// tree, a tie takes the other lane's proposal.
__global__ void k_argmin(const double *dS, const int32_t *prop, int32_t B, int32_t np,
                         int32_t *best, double *dS_best) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  double v[32];
  int32_t b[32];
#pragma unroll
  for (int l = 0; l < 32; ++l) {
    if (l < np) {
      v[l] = dS[(int64_t)r * np + l];
      b[l] = prop[(int64_t)r * np + l];
    } else {
      v[l] = 1.7976931348623157e308;
      b[l] = 0;
    }
  }
#pragma unroll
  for (int off = 16; off >= 1; off >>= 1) {
#pragma unroll
    for (int l = 0; l < 16; ++l) {
      if (l < off) {
        const double a_ = v[l], o = v[l + off];
        const double mn = (a_ != a_) ? o : (o < a_ ? o : a_);
        if (mn == o) b[l] = b[l + off];
        v[l] = mn;
      }
    }
  }
  best[r] = b[0];
  dS_best[r] = v[0];
}

void launch_argmin(const double *dS, const int32_t *prop, int32_t B, int32_t np, int32_t *best,
                   double *dS_best, hipStream_t st) {
  k_argmin<<<grid_for(B, 128), 128, 0, st>>>(dS, prop, B, np, best, dS_best);
}

// One round of pointer jumping on the merge forest par[B] (par[x] == x: a root): every
// block follows up to 64 links and keeps the ancestor reached.  Only thread x writes par[x],
// and a concurrent update only moves a pointer closer to its root, so after round k every
// pointer is the root or an ancestor at least 64^k links up.  The first round also flags
// the roots.
__global__ void k_jump(int32_t *par, int32_t B, uint32_t *is_root) {
  const int32_t x = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (x >= B) return;
  int32_t r = par[x];
  if (is_root) is_root[x] = r == x ? 1u : 0u;
  for (int k = 0; k < 64; ++k) {
    const int32_t p = par[r];
    if (p == r) break;
    r = p;
  }
  par[x] = r;
}

// lab[v] = rank of the root of lab[v]'s block (rank: exclusive scan of the root flags)
__global__ void k_relabel_roots(int32_t *lab, const int32_t *par, const uint32_t *rank,
                                int64_t N) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < N) lab[v] = (int32_t)rank[par[lab[v]]];
}

void launch_merge_relabel(int32_t *par, int32_t B, uint32_t *rank, uint32_t *tmp, int32_t *lab,
                          int64_t N, hipStream_t st) {
  int rounds = 1;
  for (int64_t reach = 64; reach < B; reach *= 64) ++rounds;
  for (int k = 0; k < rounds; ++k)
    k_jump<<<grid_for(B, 256), 256, 0, st>>>(par, B, k == 0 ? rank : nullptr);
  scan_exclusive(rank, rank, B, nullptr, tmp, st);
  k_relabel_roots<<<grid_for(N, 256), 256, 0, st>>>(lab, par, rank, N);
}

__global__ void k_iota(int32_t *x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (int32_t)i;
}


__global__ void k_order_keys(const double *x, int32_t n, uint64_t *keys, uint32_t *vals) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  keys[r] = order_key(x[r]);
  vals[r] = (uint32_t)r;
}
void launch_order_keys(const double *x, int32_t n, uint64_t *keys, uint32_t *vals,
                       hipStream_t st) {
  k_order_keys<<<grid_for(n, 256), 256, 0, st>>>(x, n, keys, vals);
}

// owner[j] = v for every entry j of vertex v's CSR row
__global__ void k_owners(const int32_t *ptr, int64_t N, int32_t *owner) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x)
    for (int32_t j = ptr[v]; j < ptr[v + 1]; ++j) owner[j] = (int32_t)v;
}
void launch_owners(const int32_t *ptr, int64_t N, int32_t *owner, hipStream_t st) {
  if (N > 0) k_owners<<<grid_for(N, 256), 256, 0, st>>>(ptr, N, owner);
}

void launch_iota(int32_t *x, int64_t n, hipStream_t st) {
  k_iota<<<grid_for(n, 256), 256, 0, st>>>(x, n);
}

}  // namespace gsap
