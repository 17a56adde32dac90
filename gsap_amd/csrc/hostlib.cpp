// Host-side helpers of the C ABI: the seeded DC-SBM benchmark generator, the
// GraphChallenge TSV reader (GSAP/load.hpp:6-133), pairwise precision / recall
// (GSAP/evaluate/evaluate.hpp:130-181) and the canonical log2 on the host.
#include <algorithm>
#include <limits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gsap/gsap.h"
#include "common.h"

using namespace gsap;

namespace {

struct Gen {
  uint32_t k0, k1;
  Philox at(uint32_t idx, uint32_t stage, uint32_t sub) const {
    return philox(idx, stage, sub, TAG_GEN, k0, k1);
  }
};

inline double unit(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }

// Gamma(alpha, 1) by Marsaglia & Tsang (2000); alpha < 1 via Gamma(alpha+1) * U^(1/alpha).
double gamma_draw(const Gen &g, uint32_t k, double alpha) {
  double boost = 1.0;
  if (alpha < 1.0) {
    Philox u = g.at(k, 1, 0xFFFFFFFFu);
    boost = std::pow(unit(u.x0), 1.0 / alpha);
    alpha += 1.0;
  }
  const double d = alpha - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
  for (uint32_t attempt = 0;; ++attempt) {
    Philox u = g.at(k, 1, attempt);
    double z = std::sqrt(-2.0 * std::log(unit(u.x0))) * std::cos(6.283185307179586 * unit(u.x1));
    double v = 1.0 + c * z;
    if (v <= 0.0) continue;
    v = v * v * v;
    double uu = unit(u.x2);
    if (std::log(uu) < 0.5 * z * z + d - d * v + d * std::log(v)) return d * v * boost;
  }
}

// first index with cdf[i] > target (cdf strictly cumulative)
int64_t pick_cdf(const uint64_t *cdf, int64_t n, uint64_t target) {
  return std::upper_bound(cdf, cdf + n, target) - cdf;
}

uint64_t mulhi64(uint32_t x, uint64_t n) {  // floor(x * n / 2^32), n < 2^32
  return ((uint64_t)x * n) >> 32;
}

struct PairSet {
  std::vector<uint64_t> t;
  uint64_t mask;
  explicit PairSet(int64_t n) {
    uint64_t cap = 1024;
    while (cap < (uint64_t)(2 * n + 16)) cap <<= 1;
    t.assign(cap, ~0ull);
    mask = cap - 1;
  }
  bool insert(uint64_t k) {
    uint64_t h = (k * 0x9E3779B97F4A7C15ull) >> 20;
    for (uint64_t i = h & mask;; i = (i + 1) & mask) {
      if (t[i] == k) return false;
      if (t[i] == ~0ull) { t[i] = k; return true; }
    }
  }
};

void build_csr(int64_t N, const std::vector<int32_t> &src, const std::vector<int32_t> &dst,
               const std::vector<int32_t> &w, int32_t *ptr, int32_t *adj, int32_t *wout) {
  // stable counting sort by src: keeps input order inside each row
  std::vector<int64_t> cnt(N + 1, 0);
  for (int32_t s : src) cnt[s + 1]++;
  for (int64_t v = 0; v < N; ++v) cnt[v + 1] += cnt[v];
  for (int64_t v = 0; v <= N; ++v) ptr[v] = (int32_t)cnt[v];
  std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
  for (size_t e = 0; e < src.size(); ++e) {
    int64_t p = cur[src[e]]++;
    adj[p] = dst[e];
    wout[p] = w[e];
  }
}

int host_threads() {
  unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(16u, h));
}

// f(t) on nthr threads, t = 0 .. nthr-1
template <class F>
void run_threads(int nthr, F &&f) {
  std::vector<std::thread> ts;
  for (int t = 1; t < nthr; ++t) ts.emplace_back([&f, t] { f(t); });
  f(0);
  for (auto &t : ts) t.join();
}

// f(lo, hi) over [0, n) split into nthr contiguous ranges
template <class F>
void parallel_for(int nthr, uint64_t n, F &&f) {
  if (nthr <= 1 || n < 4096) { f(0, n); return; }
  std::vector<std::thread> ts;
  const uint64_t per = (n + nthr - 1) / nthr;
  for (int t = 0; t < nthr; ++t) {
    const uint64_t lo = std::min(n, per * t), hi = std::min(n, lo + per);
    if (lo < hi) ts.emplace_back([&f, lo, hi] { f(lo, hi); });
  }
  for (auto &t : ts) t.join();
}

// LSD radix sort of 64-bit keys, 16-bit digits; digits that are equal in every key are
// skipped. Per-thread histograms keep the scatter stable.
void radix_sort_u64(std::vector<uint64_t> &k, std::vector<uint64_t> &tmp, int nthr) {
  const uint64_t n = k.size();
  if (n < 2) return;
  tmp.resize(n);
  uint64_t all_or = 0, all_and = ~0ull;
  for (uint64_t x : k) { all_or |= x; all_and &= x; }
  const uint64_t per = (n + nthr - 1) / nthr;
  std::vector<uint64_t> hist((size_t)nthr << 16);
  for (int sh = 0; sh < 64; sh += 16) {
    if ((((all_or ^ all_and) >> sh) & 0xFFFF) == 0) continue;
    std::fill(hist.begin(), hist.end(), 0);
    run_threads(nthr, [&](uint64_t t) {
      uint64_t *h = &hist[t << 16];
      for (uint64_t i = per * t; i < std::min(n, per * (t + 1)); ++i) h[(k[i] >> sh) & 0xFFFF]++;
    });
    uint64_t run = 0;
    for (uint32_t d = 0; d < 65536; ++d)
      for (int t = 0; t < nthr; ++t) {
        uint64_t c = hist[((size_t)t << 16) + d];
        hist[((size_t)t << 16) + d] = run;
        run += c;
      }
    run_threads(nthr, [&](uint64_t t) {
      uint64_t *h = &hist[t << 16];
      for (uint64_t i = per * t; i < std::min(n, per * (t + 1)); ++i)
        tmp[h[(k[i] >> sh) & 0xFFFF]++] = k[i];
    });
    k.swap(tmp);
  }
}

// keys sorted ascending as (major << 32 | minor), all distinct -> CSR with unit weights
void csr_from_sorted_keys(int64_t N, const std::vector<uint64_t> &k, int32_t *ptr, int32_t *adj,
                          int32_t *w) {
  std::fill(ptr, ptr + N + 1, 0);
  for (uint64_t x : k) ptr[(x >> 32) + 1]++;
  for (int64_t v = 0; v < N; ++v) ptr[v + 1] += ptr[v];
  for (size_t i = 0; i < k.size(); ++i) {
    adj[i] = (int32_t)(uint32_t)k[i];
    w[i] = 1;
  }
}

}  // namespace

extern "C" {

double gsap_log2(double x) { return clog2_host(x); }

void gsap_golden_init(gsap_golden_slot slots[3]) {
  // OldData() (graph.hpp:62-69): B = 0 and S = inf in every slot; slot k owns snapshot k
  for (int32_t k = 0; k < 3; ++k)
    slots[k] = gsap_golden_slot{0, 0, std::numeric_limits<double>::infinity(), k, 0};
}

int32_t gsap_converged(const double *dE, int32_t itr, int32_t W, double S, int32_t bracketed,
                       double threshold1, double threshold2) {
  // partition.hpp:918-934: from itr = W-1 on, the mean of the W-1 sweeps before itr (the
  // current one is outside the window) against threshold1 * S while the bracket is open,
  // threshold2 * S once all three slots hold a finite S
  if (itr < W - 1) return 0;
  double mean = 0.0;
  for (int32_t i = itr - W + 1; i < itr; ++i) mean += dE[i];
  mean /= (double)(W - 1);
  return (-mean < (bracketed ? threshold2 : threshold1) * S) ? 1 : 0;
}

void gsap_golden_step(gsap_golden_slot slots[3], const gsap_golden_slot *P, float rate,
                      int32_t *store, int32_t *restore, gsap_golden_slot *next,
                      int32_t *optimal) {
  gsap_golden_slot &large = slots[0], &med = slots[1], &small = slots[2];
  // partition.hpp:89-161: a result no worse than med takes med's place (med moving to large
  // when it had more blocks, else to small, the two snapshot buffers exchanged); a worse one
  // replaces small or large
  gsap_golden_slot *dst;
  if (P->S <= med.S) {
    std::swap(med.B > P->B ? large : small, med);
    dst = &med;
  } else {
    dst = med.B > P->B ? &small : &large;
  }
  dst->B = P->B;
  dst->B_to_merge = P->B_to_merge;
  dst->S = P->S;
  *store = dst->snap;
  // partition.hpp:163-320: the slot restored and the next number of merges
  const gsap_golden_slot *from = &med;
  int32_t Bm, opt = 0;
  if (std::isinf(small.S)) {
    Bm = (int32_t)((float)med.B * rate);  // (int)P.B * rate: a float product, truncated
    opt = Bm == 0;
  } else if (large.B - small.B == 2) {
    opt = 1;
    Bm = med.B_to_merge;
  } else if (large.B - med.B >= med.B - small.B) {
    from = &large;
    Bm = (large.B - med.B) - (int32_t)std::round((double)(large.B - med.B) * 0.618);
  } else {
    Bm = (med.B - small.B) - (int32_t)std::round((double)(med.B - small.B) * 0.618);
  }
  // not in the reference: a bracket step without merges (two slots of equal B, possible
  // when a merge phase runs out of candidates) ends the search instead of repeating a level
  if (Bm <= 0) opt = 1;
  *restore = from->snap;
  *next = gsap_golden_slot{from->B, Bm, from->S, from->snap, 0};
  *optimal = opt;
}

void gsap_log2_tables(double *inv_c, double *log2_c, double *poly) {
  std::memcpy(inv_c, h_log2_inv, sizeof h_log2_inv);
  std::memcpy(log2_c, h_log2_lg, sizeof h_log2_lg);
  std::memcpy(poly, h_log2_poly, sizeof h_log2_poly);
}

void gsap_log2f_n(const float *x, float *y, int64_t n) {
  for (int64_t k = 0; k < n; ++k) y[k] = flog2_host(x[k]);
}

int gsap_generate_dcsbm_i32(const gsap_dcsbm_params *p, int64_t *E_out, int32_t *out_ptr,
                            int32_t *out_adj, int32_t *out_w, int32_t *in_ptr, int32_t *in_adj,
                            int32_t *in_w, int32_t *truth) {
  if (!p || !E_out) return GSAP_ERR_INVALID;
  const int64_t N = p->N;
  const int32_t K = p->K;
  if (N <= 1 || N >= ((int64_t)1 << 31) || K < 1 || K > N || p->mean_deg <= 0 || p->rho < 0 ||
      p->deg_min < 1 || p->deg_max < p->deg_min)
    return GSAP_ERR_INVALID;
  if ((uint64_t)N * (uint64_t)p->deg_max >= ((uint64_t)1 << 32)) return GSAP_ERR_INVALID;
  const int64_t E = std::llround((double)N * p->mean_deg);
  if (E <= 0 || E > N * (N - 1) / 4 || E >= ((int64_t)1 << 30)) return GSAP_ERR_INVALID;
  *E_out = E;
  if (!out_ptr) return GSAP_OK;
  if (!out_adj || !out_w || !in_ptr || !in_adj || !in_w) return GSAP_ERR_INVALID;
  const Gen g{(uint32_t)p->seed, (uint32_t)(p->seed >> 32)};

  // 1. block sizes: near-equal, or Dirichlet(alpha) with every block non-empty
  std::vector<int64_t> size(K);
  if (p->alpha <= 0) {
    for (int32_t k = 0; k < K; ++k) size[k] = N / K + (k < N % K ? 1 : 0);
  } else {
    std::vector<double> gk(K);
    double tot = 0;
    for (int32_t k = 0; k < K; ++k) { gk[k] = gamma_draw(g, (uint32_t)k, p->alpha); tot += gk[k]; }
    const int64_t spare = N - K;
    std::vector<double> frac(K);
    int64_t used = 0;
    for (int32_t k = 0; k < K; ++k) {
      double x = gk[k] / tot * (double)spare;
      size[k] = 1 + (int64_t)std::floor(x);
      frac[k] = x - std::floor(x);
      used += size[k];
    }
    std::vector<int32_t> ord(K);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return frac[a] > frac[b]; });
    for (int64_t i = 0; used < N; ++i, ++used) size[ord[i % K]]++;
  }
  // 2. vertex -> block, shuffled (Fisher-Yates)
  std::vector<int32_t> blk(N);
  {
    int64_t v = 0;
    for (int32_t k = 0; k < K; ++k)
      for (int64_t i = 0; i < size[k]; ++i) blk[v++] = k;
    for (int64_t i = N - 1; i > 0; --i) {
      Philox u = g.at((uint32_t)i, 2, 0);
      int64_t j = (int64_t)mulhi64(u.x0, (uint64_t)(i + 1));
      std::swap(blk[i], blk[j]);
    }
  }
  // 3. degree propensities: discrete power law k^-exponent on [deg_min, deg_max]
  const int32_t nk = p->deg_max - p->deg_min + 1;
  std::vector<double> pcdf(nk);
  {
    double s = 0;
    for (int32_t i = 0; i < nk; ++i) { s += std::pow((double)(p->deg_min + i), -p->exponent); pcdf[i] = s; }
    for (int32_t i = 0; i < nk; ++i) pcdf[i] /= s;
  }
  auto theta = [&](uint32_t x) {
    double u = unit(x);
    int32_t i = (int32_t)(std::upper_bound(pcdf.begin(), pcdf.end(), u) - pcdf.begin());
    return (uint64_t)(p->deg_min + std::min(i, nk - 1));
  };
  std::vector<uint64_t> cdf_out(N);
  std::vector<std::vector<int32_t>> members(K);
  std::vector<std::vector<uint64_t>> cdf_in(K);
  {
    uint64_t acc = 0;
    for (int64_t v = 0; v < N; ++v) {
      Philox u = g.at((uint32_t)v, 3, 0);
      acc += theta(u.x0);
      cdf_out[v] = acc;
      members[blk[v]].push_back((int32_t)v);
      uint64_t prev = cdf_in[blk[v]].empty() ? 0 : cdf_in[blk[v]].back();
      cdf_in[blk[v]].push_back(prev + theta(u.x1));
    }
  }
  const uint64_t tot_out = cdf_out[N - 1];
  const uint32_t thr = (uint32_t)std::min(4294967295.0, std::floor(p->rho / (1.0 + p->rho) * 4294967296.0));
  // 4. edges: source ~ theta_out, destination block intra w.p. rho/(1+rho) else uniform
  //    over the other blocks, destination ~ theta_in inside it; no self-loops, no duplicates.
  //    Attempt a depends only on (seed, a), so candidates are drawn in parallel chunks and
  //    then accepted in attempt order: the first E distinct non-loop pairs.
  auto candidate = [&](uint64_t a) -> uint64_t {
    Philox u = g.at((uint32_t)a, 4, (uint32_t)(a >> 32));
    int64_t s = pick_cdf(cdf_out.data(), N, mulhi64(u.x0, tot_out));
    int32_t bs = blk[s], bd = bs;
    if (K > 1 && u.x1 >= thr) {
      bd = (int32_t)mulhi64(u.x2, (uint64_t)(K - 1));
      if (bd >= bs) ++bd;
    }
    const std::vector<uint64_t> &ci = cdf_in[bd];
    int64_t t = members[bd][pick_cdf(ci.data(), (int64_t)ci.size(), mulhi64(u.x3, ci.back()))];
    return t == s ? ~0ull : (((uint64_t)s << 32) | (uint64_t)t);
  };
  const int nthr = host_threads();
  std::vector<uint64_t> keys;
  keys.reserve(E);
  {
    PairSet seen(E);
    const uint64_t max_attempts = (uint64_t)E * 200 + 100000;
    const uint64_t chunk = std::max<uint64_t>(1 << 16, (uint64_t)E / 8);
    std::vector<uint64_t> cand(chunk);
    for (uint64_t a0 = 0; (int64_t)keys.size() < E; a0 += chunk) {
      if (a0 >= max_attempts) return GSAP_ERR_INVALID;
      parallel_for(nthr, chunk, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) cand[i] = candidate(a0 + i);
      });
      for (uint64_t i = 0; i < chunk && (int64_t)keys.size() < E; ++i) {
        if (a0 + i >= max_attempts) return GSAP_ERR_INVALID;
        if (cand[i] == ~0ull || !seen.insert(cand[i])) continue;
        keys.push_back(cand[i]);
      }
    }
  }
  // 5. CSR views, each row ordered by neighbour id
  {
    std::vector<uint64_t> tmp(E);
    radix_sort_u64(keys, tmp, nthr);
    csr_from_sorted_keys(N, keys, out_ptr, out_adj, out_w);
    for (int64_t i = 0; i < E; ++i) keys[i] = (keys[i] << 32) | (keys[i] >> 32);
    radix_sort_u64(keys, tmp, nthr);
    csr_from_sorted_keys(N, keys, in_ptr, in_adj, in_w);
  }
  if (truth) std::copy(blk.begin(), blk.end(), truth);
  return GSAP_OK;
}

int gsap_read_tsv_i32(const char *path, int64_t *N_out, int64_t *E_out, int32_t *out_ptr,
                      int32_t *out_adj, int32_t *out_w, int32_t *in_ptr, int32_t *in_adj,
                      int32_t *in_w) {
  if (!path || !N_out || !E_out) return GSAP_ERR_INVALID;
  std::ifstream f(path);
  if (!f.is_open()) return GSAP_ERR_IO;
  std::vector<int32_t> src, dst, w;
  std::string line;
  int64_t N = 0;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    long long a, b;
    double x = 1.0;
    if (!(is >> a)) continue;  // blank line
    if (!(is >> b)) return GSAP_ERR_IO;
    if (!(is >> x)) x = 1.0;
    if (a < 1 || b < 1 || a >= ((long long)1 << 31) || b >= ((long long)1 << 31)) return GSAP_ERR_IO;
    src.push_back((int32_t)(a - 1));
    dst.push_back((int32_t)(b - 1));
    w.push_back((int32_t)x);  // load.hpp:72-73: stof then integer cast
    N = std::max<int64_t>(N, std::max(a, b));
  }
  *N_out = N;
  *E_out = (int64_t)src.size();
  if (!out_ptr) return GSAP_OK;
  if (!out_adj || !out_w || !in_ptr || !in_adj || !in_w) return GSAP_ERR_INVALID;
  build_csr(N, src, dst, w, out_ptr, out_adj, out_w);
  build_csr(N, dst, src, w, in_ptr, in_adj, in_w);
  return GSAP_OK;
}

int gsap_pairwise_i32(int64_t N, const int32_t *truth, const int32_t *given, double *precision,
                      double *recall) {
  if (N <= 0 || !truth || !given || !precision || !recall) return GSAP_ERR_INVALID;
  int32_t Bt = 0, Bg = 0;
  for (int64_t i = 0; i < N; ++i) {
    if (given[i] < 0) return GSAP_ERR_INVALID;
    Bt = std::max(Bt, truth[i] + 1);
    Bg = std::max(Bg, given[i] + 1);
  }
  std::vector<int64_t> a(Bt, 0), b(Bg, 0);
  std::vector<uint64_t> pairs;
  pairs.reserve(N);
  for (int64_t i = 0; i < N; ++i) {
    if (truth[i] < 0) continue;  // unknown truth, evaluate.hpp:43-48
    a[truth[i]]++;
    b[given[i]]++;
    pairs.push_back(((uint64_t)truth[i] << 32) | (uint32_t)given[i]);
  }
  std::sort(pairs.begin(), pairs.end());
  int64_t same = 0;
  for (size_t i = 0; i < pairs.size();) {
    size_t j = i;
    while (j < pairs.size() && pairs[j] == pairs[i]) ++j;
    int64_t n = (int64_t)(j - i);
    same += n * (n - 1) / 2;
    i = j;
  }
  int64_t sa = 0, sb = 0;
  for (int64_t x : a) sa += x * (x - 1) / 2;
  for (int64_t x : b) sb += x * (x - 1) / 2;
  *recall = sa ? (double)same / (double)sa : 0.0;
  *precision = sb ? (double)same / (double)sb : 0.0;
  return GSAP_OK;
}


// Full quality report of a partition against a truth labelling (the metrics of
// bf::evaluate, GSAP/evaluate/evaluate.hpp:19-265, with int64 pair counts and FP64 sums).
// Vertices whose truth is negative are left out (evaluate.hpp:43-48).  The accuracy is the
// largest total of a one-to-one matching between truth and given blocks in the contingency
// table (Kuhn-Munkres, evaluate.hpp:69-96), divided by the number of vertices; it is
// computed when the smaller side times the larger side stays within 64M cells.
int gsap_quality_i32(int64_t N, const int32_t *truth, const int32_t *given, gsap_quality *q) {
  if (N <= 0 || !truth || !given || !q) return GSAP_ERR_INVALID;
  int32_t Bt = 0, Bg = 0;
  for (int64_t i = 0; i < N; ++i) {
    if (given[i] < 0) return GSAP_ERR_INVALID;
    Bt = std::max(Bt, truth[i] + 1);
    Bg = std::max(Bg, given[i] + 1);
  }
  // contingency table as sorted (truth, given) pairs with counts
  std::vector<uint64_t> key;
  key.reserve(N);
  for (int64_t i = 0; i < N; ++i)
    if (truth[i] >= 0) key.push_back(((uint64_t)truth[i] << 32) | (uint32_t)given[i]);
  std::sort(key.begin(), key.end());
  const int64_t n = (int64_t)key.size();
  struct Cell { int32_t t, g; int64_t c; };
  std::vector<Cell> cells;
  for (size_t i = 0; i < key.size();) {
    size_t j = i;
    while (j < key.size() && key[j] == key[i]) ++j;
    cells.push_back(Cell{(int32_t)(key[i] >> 32), (int32_t)(uint32_t)key[i], (int64_t)(j - i)});
    i = j;
  }
  std::vector<int64_t> rs(std::max(Bt, 1), 0), cs(std::max(Bg, 1), 0);
  for (const Cell &c : cells) { rs[c.t] += c.c; cs[c.g] += c.c; }
  std::memset(q, 0, sizeof(*q));
  q->n = n;
  int32_t nonempty_t = 0;
  for (int64_t x : rs) nonempty_t += x > 0;
  q->B_truth = nonempty_t;
  q->B_given = Bg;
  if (n == 0) return GSAP_ERR_INVALID;
  // pair counts
  auto ch2 = [](int64_t x) { return x * (x - 1) / 2; };
  int64_t same = 0, sa = 0, sb = 0;
  for (const Cell &c : cells) same += ch2(c.c);
  for (int64_t x : rs) sa += ch2(x);
  for (int64_t x : cs) sb += ch2(x);
  const double pairs = (double)ch2(n);
  q->recall = sa ? (double)same / (double)sa : 0.0;
  q->precision = sb ? (double)same / (double)sb : 0.0;
  // agreements: pairs together in both + pairs apart in both
  const double apart_both = pairs - (double)sa - (double)sb + (double)same;
  q->rand_index = pairs > 0 ? ((double)same + apart_both) / pairs : 1.0;
  const double expct = pairs > 0 ? (double)sa * (double)sb / pairs : 0.0;
  const double den = 0.5 * ((double)sa + (double)sb) - expct;
  q->adjusted_rand_index = den != 0.0 ? ((double)same - expct) / den : 1.0;
  // information-theoretic: entropies and mutual information in nats
  double H1 = 0.0, H2 = 0.0, MI = 0.0;
  const double dn = (double)n;
  for (int64_t x : rs) if (x > 0) { const double p = (double)x / dn; H1 -= p * std::log(p); }
  for (int64_t x : cs) if (x > 0) { const double p = (double)x / dn; H2 -= p * std::log(p); }
  for (const Cell &c : cells) {
    const double pij = (double)c.c / dn;
    MI += pij * std::log(pij / (((double)rs[c.t] / dn) * ((double)cs[c.g] / dn)));
  }
  q->nmi = 2.0 * std::fabs(MI) / (std::fabs(H1) + std::fabs(H2));
  // optimal one-to-one matching (Hungarian algorithm with row and column potentials on the
  // rectangular table, rows = the smaller side; cost = -count, so the minimum-cost
  // assignment maximises the matched count)
  const bool tr = Bt > Bg;
  const int64_t R = tr ? Bg : Bt, Cn = tr ? Bt : Bg;
  if (R >= 1 && R * Cn <= ((int64_t)64 << 20)) {
    std::vector<int64_t> cost((size_t)(R + 1) * (Cn + 1), 0);
    for (const Cell &c : cells) {
      const int64_t r = tr ? c.g : c.t, k = tr ? c.t : c.g;
      cost[(size_t)(r + 1) * (Cn + 1) + (k + 1)] = -c.c;
    }
    const int64_t INF = std::numeric_limits<int64_t>::max() / 4;
    std::vector<int64_t> u(R + 1, 0), v(Cn + 1, 0), minv(Cn + 1);
    std::vector<int64_t> owner(Cn + 1, 0), prev(Cn + 1, 0);  // column -> row, path links
    std::vector<char> seen(Cn + 1);
    for (int64_t r = 1; r <= R; ++r) {
      owner[0] = r;
      int64_t col = 0;
      std::fill(minv.begin(), minv.end(), INF);
      std::fill(seen.begin(), seen.end(), 0);
      do {
        seen[col] = 1;
        const int64_t row = owner[col];
        int64_t delta = INF, next = 0;
        const int64_t *crow = &cost[(size_t)row * (Cn + 1)];
        for (int64_t k = 1; k <= Cn; ++k) {
          if (seen[k]) continue;
          const int64_t red = crow[k] - u[row] - v[k];
          if (red < minv[k]) { minv[k] = red; prev[k] = col; }
          if (minv[k] < delta) { delta = minv[k]; next = k; }
        }
        for (int64_t k = 0; k <= Cn; ++k) {
          if (seen[k]) { u[owner[k]] += delta; v[k] -= delta; }
          else minv[k] -= delta;
        }
        col = next;
      } while (owner[col] != 0);
      do {  // flip the augmenting path
        const int64_t back = prev[col];
        owner[col] = owner[back];
        col = back;
      } while (col != 0);
    }
    int64_t matched = 0;
    for (int64_t k = 1; k <= Cn; ++k)
      if (owner[k] != 0) matched -= cost[(size_t)owner[k] * (Cn + 1) + k];
    q->accuracy = (double)matched / dn;
    q->accuracy_computed = 1;
  }
  return GSAP_OK;
}


// ---- binary CSR cache: "GSAPCSR1", flags (bit 0: truth present), N, E, then int32 arrays
// out_ptr[N+1], out_adj[E], out_w[E], in_ptr[N+1], in_adj[E], in_w[E], [truth[N]], then a
// 64-bit checksum of the arrays (little-endian host layout).
namespace {
const char kCsrMagic[8] = {'G', 'S', 'A', 'P', 'C', 'S', 'R', '1'};

uint64_t csr_checksum(uint64_t h, const int32_t *p, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    h ^= (uint32_t)p[i];
    h *= 0x100000001B3ull;
    h ^= h >> 29;
  }
  return h;
}
}  // namespace

int gsap_save_csr_i32(const char *path, const gsap_graph_i32 *g, const int32_t *truth) {
  if (!path || !g || g->N <= 0 || g->E < 0 || !g->out_ptr || !g->in_ptr) return GSAP_ERR_INVALID;
  FILE *f = std::fopen(path, "wb");
  if (!f) return GSAP_ERR_IO;
  const uint32_t flags = truth ? 1u : 0u, pad = 0;
  const int64_t N = g->N, E = g->E;
  const int32_t *arr[7] = {g->out_ptr, g->out_adj, g->out_w, g->in_ptr, g->in_adj, g->in_w, truth};
  const int64_t len[7] = {N + 1, E, E, N + 1, E, E, N};
  bool ok = std::fwrite(kCsrMagic, 1, 8, f) == 8 && std::fwrite(&flags, 4, 1, f) == 1 &&
            std::fwrite(&pad, 4, 1, f) == 1 && std::fwrite(&N, 8, 1, f) == 1 &&
            std::fwrite(&E, 8, 1, f) == 1;
  uint64_t h = 0xCBF29CE484222325ull;
  for (int k = 0; k < 7 && ok; ++k) {
    if (k == 6 && !truth) break;
    if (len[k] > 0 && !arr[k]) { ok = false; break; }
    ok = std::fwrite(arr[k], 4, (size_t)len[k], f) == (size_t)len[k];
    h = csr_checksum(h, arr[k], len[k]);
  }
  ok = ok && std::fwrite(&h, 8, 1, f) == 1;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? GSAP_OK : GSAP_ERR_IO;
}

int gsap_load_csr_i32(const char *path, int64_t *N_out, int64_t *E_out, int32_t *has_truth,
                      int32_t *out_ptr, int32_t *out_adj, int32_t *out_w, int32_t *in_ptr,
                      int32_t *in_adj, int32_t *in_w, int32_t *truth) {
  if (!path || !N_out || !E_out || !has_truth) return GSAP_ERR_INVALID;
  FILE *f = std::fopen(path, "rb");
  if (!f) return GSAP_ERR_IO;
  char magic[8];
  uint32_t flags = 0, pad = 0;
  int64_t N = 0, E = 0;
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kCsrMagic, 8) == 0 &&
            std::fread(&flags, 4, 1, f) == 1 && std::fread(&pad, 4, 1, f) == 1 &&
            std::fread(&N, 8, 1, f) == 1 && std::fread(&E, 8, 1, f) == 1 && N > 0 && E >= 0 &&
            N < ((int64_t)1 << 31) && E < ((int64_t)1 << 31);
  if (!ok) { std::fclose(f); return GSAP_ERR_IO; }
  *N_out = N;
  *E_out = E;
  *has_truth = (flags & 1u) ? 1 : 0;
  if (!out_ptr) { std::fclose(f); return GSAP_OK; }  // sizes only
  if (!out_adj || !out_w || !in_ptr || !in_adj || !in_w || (*has_truth && !truth)) {
    std::fclose(f);
    return GSAP_ERR_INVALID;
  }
  int32_t *arr[7] = {out_ptr, out_adj, out_w, in_ptr, in_adj, in_w, truth};
  const int64_t len[7] = {N + 1, E, E, N + 1, E, E, N};
  uint64_t h = 0xCBF29CE484222325ull, stored = 0;
  for (int k = 0; k < 7 && ok; ++k) {
    if (k == 6 && !*has_truth) break;
    ok = std::fread(arr[k], 4, (size_t)len[k], f) == (size_t)len[k];
    if (ok) h = csr_checksum(h, arr[k], len[k]);
  }
  ok = ok && std::fread(&stored, 8, 1, f) == 1 && stored == h;
  std::fclose(f);
  return ok ? GSAP_OK : GSAP_ERR_IO;
}

}  // extern "C"
