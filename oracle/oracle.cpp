// CPU oracle for the GSAP stochastic-block-partitioning hot path.
//
// TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
// Nothing in gsap_amd/ includes, links or calls this file.
//
// This is a plain, single-threaded restatement of the reference algorithm:
//   GSAP/partition.hpp:852-954   partition(): merge -> nodal sweeps -> MDL -> golden section
//   GSAP/partition.hpp:324-454   propose_block_merge
//   GSAP/partition.hpp:26-78     carry_out_best_merges
//   GSAP/partition.hpp:456-662   propose_nodal_move
//   GSAP/partition.hpp:665-704   compute_overall_entropy_cuda
//   GSAP/partition.hpp:707-849   calculate_new_csr
//   GSAP/partition.hpp:80-322    prepare_for_partition_next
//   GSAP/kernels.hpp             every proposal / delta-S formula
//
// Decisions on the reference's quirks (numbers follow the quirk register):
//   Q1-Q3  clock-seeded curand + float prefix sampling -> counter-based Philox4x32-10 with
//          integer sampling of the same distribution (mulhi(u32, n), integer prefix walk,
//          exact integer coin u*(deg+B) < B*2^32).  Q3's "skip r, carry its mass forward"
//          and the block-0 fall-through are kept.
//   Q4-Q8  kept (self-proposals, per-block coin/sample, greedy accept dS<0, batch-Jacobi,
//          4 contiguous batches of floor(N/4) with the tail never moving).
//   Q9     fixed: the s-side old column term excludes {r, s} (not s's own i-th proposal).
//   Q10,Q11 kept: merged row keeps columns r and s apart; new rows use old column degrees.
//   Q12,Q13 fixed: node-only entries are run-merged and use the node's own weights.
//   Q14,Q15,Q18,Q19 fixed (memory/size bugs; every one of the np*B proposals is computed).
//   Q16    kept in form (model term E(1+x)ln(1+x) - x ln x + N ln B) but evaluated in FP64.
//   Q17    fixed: degrees and edge counts are kept apart; weights travel with their edges.
//   Q20    argmin follows the reference's 32-lane shuffle-down tree exactly.
//   Q21    kept: emptied blocks keep their ids, B is unchanged by nodal moves.
//   Q22    union-find with identical semantics; the loop also stops after B candidates.
//   Q25    pair counts in int64.
// All delta-S / MDL arithmetic is FP64 with a canonical log2 and a canonical summation
// order (64 strided partials + binary tree), which the HIP kernels reproduce exactly.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>
#include <vector>

namespace {

// ------------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11.
constexpr uint32_t PH_M0 = 0xD2511F53u, PH_M1 = 0xCD9E8D57u;
constexpr uint32_t PH_W0 = 0x9E3779B9u, PH_W1 = 0xBB67AE85u;

void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
            uint32_t out[4]) {
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
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Counter word 3 tags the purpose of a draw.
constexpr uint32_t TAG_MERGE = 0x4D455247u;  // block-merge proposal (r, i)
constexpr uint32_t TAG_NODE = 0x4E4F4445u;   // nodal proposal of a vertex
constexpr uint32_t TAG_BLOCK = 0x424C4B53u;  // nodal per-block sample + coin

struct Rng {
  uint32_t k0, k1, chain;
  void draw(uint32_t idx, uint32_t phase, uint32_t tag, uint32_t out[4]) const {
    philox(idx, phase, chain, tag, k0, k1, out);
  }
};

inline uint32_t mulhi(uint32_t x, uint64_t n) { return (uint32_t)(((uint64_t)x * n) >> 32); }
// random-vs-neighbour coin: true with probability B/(deg+B)   (kernels.hpp:720, 847, 1024)
inline bool coin_random(uint32_t x, int64_t deg, int64_t B) {
  return (uint64_t)x * (uint64_t)(deg + B) < ((uint64_t)B << 32);
}

// -------------------------------------------------------------- canonical log2
// log2(x) = e + log2(c_j) + P(r),  x = 2^e * m,  c_j = 1 + j/128 (top 7 mantissa bits),
// r = fma(m, 1/c_j, -1),  P(r) = sum_{k=1..9} (-1)^(k+1) r^k / (k ln 2)  (Horner, fma).
// Constants are computed here in binary128 and rounded once to double.
struct Log2Tables {
  double inv[128], lg[128], poly[9];
  Log2Tables() {
    typedef __float128 Q;
    auto qln = [](Q c) {  // ln(c) = 2 atanh((c-1)/(c+1))
      Q z = (c - 1) / (c + 1), z2 = z * z, t = z, s = 0;
      for (int k = 0; k < 200; ++k) { s += t / (Q)(2 * k + 1); t *= z2; }
      return 2 * s;
    };
    Q ln2 = qln((Q)2);
    for (int j = 0; j < 128; ++j) {
      double c = 1.0 + (double)j / 128.0;
      inv[j] = 1.0 / c;  // IEEE division: correctly rounded
      lg[j] = (double)(qln((Q)c) / ln2);
    }
    for (int k = 1; k <= 9; ++k) {
      Q a = (Q)1 / ((Q)k * ln2);
      poly[k - 1] = (double)((k & 1) ? a : -a);
    }
  }
};
const Log2Tables &tables() {
  static Log2Tables t;
  return t;
}

double clog2(double x) {
  if (x == 0.0) return -std::numeric_limits<double>::infinity();
  if (!(x > 0.0)) return std::numeric_limits<double>::quiet_NaN();
  if (std::isinf(x)) return x;
  const Log2Tables &T = tables();
  uint64_t bits;
  std::memcpy(&bits, &x, 8);
  int e = (int)((bits >> 52) & 0x7ff);
  if (e == 0) {  // subnormal: scale into the normal range
    double y = x * 18014398509481984.0;  // 2^54
    std::memcpy(&bits, &y, 8);
    e = (int)((bits >> 52) & 0x7ff) - 54;
  }
  e -= 1023;
  int j = (int)((bits >> 45) & 127);
  uint64_t mb = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
  double m;
  std::memcpy(&m, &mb, 8);
  double r = std::fma(m, T.inv[j], -1.0);
  double p = T.poly[8];
  for (int k = 7; k >= 0; --k) p = std::fma(p, r, T.poly[k]);
  p = p * r;
  return ((double)e + T.lg[j]) + p;
}

// x*log2(x)-style MDL term, same grouping as kernels.hpp:176-177:
//   p = log w - (log a + log b);  dS += w * p
inline double term(int64_t w, double la, double lb) {
  double dw = (double)w;
  return dw * (clog2(dw) - (la + lb));
}

// ------------------------------------------------------------- canonical sums
struct CSum {
  double p[64];
  CSum() { for (int i = 0; i < 64; ++i) p[i] = 0.0; }
  void add(int64_t k, double t) { p[k & 63] += t; }
  double result() {
    for (int off = 32; off >= 1; off >>= 1)
      for (int l = 0; l < off; ++l) p[l] += p[l + off];
    return p[0];
  }
};

double csum_array(const double *x, int64_t n) {
  const int64_t T = 1024;
  std::vector<double> tiles;
  for (int64_t t0 = 0; t0 < n; t0 += T) {
    CSum c;
    int64_t t1 = std::min(n, t0 + T);
    for (int64_t i = t0; i < t1; ++i) c.add(i - t0, x[i]);
    tiles.push_back(c.result());
  }
  CSum c;
  for (size_t i = 0; i < tiles.size(); ++i) c.add((int64_t)i, tiles[i]);
  return c.result();
}

// ------------------------------------------------------------------- graph
struct G {
  int64_t N, E;
  const int32_t *optr, *oadj, *ow, *iptr, *iadj, *iw;
  std::vector<int64_t> kout, kin;
  explicit G(const orc_graph *g)
      : N(g->N), E(g->E), optr(g->out_ptr), oadj(g->out_adj), ow(g->out_w), iptr(g->in_ptr),
        iadj(g->in_adj), iw(g->in_w), kout(g->N, 0), kin(g->N, 0) {
    for (int64_t v = 0; v < N; ++v) {
      for (int64_t j = optr[v]; j < optr[v + 1]; ++j) kout[v] += ow[j];
      for (int64_t j = iptr[v]; j < iptr[v + 1]; ++j) kin[v] += iw[j];
    }
  }
};

// Block matrix M[r][s] = sum of w(u->v) with b(u)=r, b(v)=s   (partition.hpp:707-849)
struct BM {
  int32_t B = 0;
  int64_t nnz = 0;
  std::vector<int64_t> rowptr, colptr, val, cval, dout, din;
  std::vector<int32_t> col, row;
  std::vector<double> ldout, ldin;
  // M[r][t] by binary search in row r (0 if absent)
  int64_t row_get(int32_t r, int32_t t) const {
    auto b = col.begin() + rowptr[r], e = col.begin() + rowptr[r + 1];
    auto it = std::lower_bound(b, e, t);
    return (it != e && *it == t) ? val[it - col.begin()] : 0;
  }
  int64_t col_get(int32_t s, int32_t t) const {
    auto b = row.begin() + colptr[s], e = row.begin() + colptr[s + 1];
    auto it = std::lower_bound(b, e, t);
    return (it != e && *it == t) ? cval[it - row.begin()] : 0;
  }
};

// f(t, lo, hi) on nthr threads over [0, n) in contiguous ranges (results are identical for
// any thread count: every index is computed by the same code, independently)
template <class F>
void par_ranges(int nthr, int64_t n, F &&f) {
  if (nthr <= 1 || n < 2 * nthr) {
    f(0, (int64_t)0, n);
    return;
  }
  std::vector<std::thread> ts;
  for (int t = 0; t < nthr; ++t) {
    const int64_t lo = n * t / nthr, hi = n * (t + 1) / nthr;
    ts.emplace_back([&f, t, lo, hi] { f(t, lo, hi); });
  }
  for (auto &t : ts) t.join();
}

void build_bm(int64_t N, const int32_t *optr, const int32_t *oadj, const int32_t *ow,
              const int32_t *lab, int32_t B, BM &m, int nthr = 1) {
  m.B = B;
  // bucket edges by source block (stable), then sort each row by destination block
  std::vector<int64_t> cnt(B + 1, 0);
  for (int64_t v = 0; v < N; ++v) cnt[lab[v] + 1] += optr[v + 1] - optr[v];
  for (int32_t r = 0; r < B; ++r) cnt[r + 1] += cnt[r];
  int64_t E = cnt[B];
  std::vector<std::pair<int32_t, int32_t>> ent(E);
  std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
  for (int64_t v = 0; v < N; ++v)
    for (int64_t j = optr[v]; j < optr[v + 1]; ++j)
      ent[cur[lab[v]]++] = std::make_pair(lab[oadj[j]], ow[j]);
  m.rowptr.assign(B + 1, 0);
  m.col.clear(); m.val.clear();
  m.dout.assign(B, 0); m.din.assign(B, 0);
  par_ranges(nthr, B, [&](int, int64_t r0, int64_t r1) {
    for (int64_t r = r0; r < r1; ++r)
      std::sort(ent.begin() + cnt[r], ent.begin() + cnt[r + 1],
                [](const std::pair<int32_t, int32_t> &x, const std::pair<int32_t, int32_t> &y) {
                  return x.first < y.first;
                });
  });
  for (int32_t r = 0; r < B; ++r) {
    auto b = ent.begin() + cnt[r], e = ent.begin() + cnt[r + 1];
    for (auto it = b; it != e;) {
      int32_t t = it->first;
      int64_t s = 0;
      while (it != e && it->first == t) { s += it->second; ++it; }
      if (s != 0) { m.col.push_back(t); m.val.push_back(s); m.dout[r] += s; m.din[t] += s; }
    }
    m.rowptr[r + 1] = (int64_t)m.col.size();
  }
  m.nnz = (int64_t)m.col.size();
  // CSC by a stable counting transpose (rows stay ascending inside each column)
  m.colptr.assign(B + 1, 0);
  for (int64_t i = 0; i < m.nnz; ++i) m.colptr[m.col[i] + 1]++;
  for (int32_t s = 0; s < B; ++s) m.colptr[s + 1] += m.colptr[s];
  m.row.assign(m.nnz, 0); m.cval.assign(m.nnz, 0);
  std::vector<int64_t> cc(m.colptr.begin(), m.colptr.end() - 1);
  for (int32_t r = 0; r < B; ++r)
    for (int64_t i = m.rowptr[r]; i < m.rowptr[r + 1]; ++i) {
      int64_t p = cc[m.col[i]]++;
      m.row[p] = r; m.cval[p] = m.val[i];
    }
  m.ldout.resize(B); m.ldin.resize(B);
  for (int32_t x = 0; x < B; ++x) { m.ldout[x] = clog2((double)m.dout[x]); m.ldin[x] = clog2((double)m.din[x]); }
}

// Old outgoing-row term of every block (kernels.hpp:143-182).
std::vector<double> row_terms(const BM &m) {
  std::vector<double> out(m.B);
  for (int32_t x = 0; x < m.B; ++x) {
    CSum c;
    for (int64_t k = m.rowptr[x]; k < m.rowptr[x + 1]; ++k)
      c.add(k - m.rowptr[x], term(m.val[k], m.ldout[x], m.ldin[m.col[k]]));
    out[x] = c.result();
  }
  return out;
}

// Weighted pick over row x (out entries) then column x (in entries): first entry whose
// running weight exceeds target, skipping `skip` (its mass carries forward).  Returns
// `fallback` when nothing qualifies.  (kernels.hpp:767-816, 978-995)
int32_t walk_block(const BM &m, int32_t x, uint64_t target, int32_t skip, int32_t fallback) {
  uint64_t cum = 0;
  for (int64_t k = m.rowptr[x]; k < m.rowptr[x + 1]; ++k) {
    cum += (uint64_t)m.val[k];
    if (cum > target && m.col[k] != skip) return m.col[k];
  }
  for (int64_t k = m.colptr[x]; k < m.colptr[x + 1]; ++k) {
    cum += (uint64_t)m.cval[k];
    if (cum > target && m.row[k] != skip) return m.row[k];
  }
  return fallback;
}

// ------------------------------------------------------------------ merge phase
// Proposal for (r, i)   (kernels.hpp:726-856)
int32_t merge_propose(const BM &m, const Rng &rng, uint32_t phase, int32_t r, int np, int i) {
  uint32_t x[4];
  rng.draw((uint32_t)((int64_t)r * np + i), phase, TAG_MERGE, x);
  int64_t B = m.B;
  int64_t deg = m.dout[r] + m.din[r];
  if (deg == 0) return (int32_t)mulhi(x[0], B);
  int32_t u = walk_block(m, r, mulhi(x[1], deg), -1, 0);
  int64_t degu = m.dout[u] + m.din[u];
  if (coin_random(x[3], degu, B)) return (int32_t)mulhi(x[0], B);
  return walk_block(m, u, mulhi(x[2], degu), r, 0);
}

// delta-S of merging r into s (kernels.hpp:143-449 with Q9 fixed, Q10 kept).  Exact form:
// the merged row and column hold t outside {r, s}, and the four entries of rows and columns
// r, s fold into one, M_rr + M_rs + M_sr + M_ss, with the merged row and column degrees; a
// proposal of r itself changes nothing and gets the largest double (never preferred).
double merge_dS(const BM &m, const std::vector<double> &OUT, int32_t r, int32_t s, bool exact) {
  if (exact && s == r) return std::numeric_limits<double>::max();
  // old column terms excluding t in {r, s}
  auto col_old = [&](int32_t c) {
    CSum cs;
    for (int64_t k = m.colptr[c]; k < m.colptr[c + 1]; ++k) {
      int32_t t = m.row[k];
      if (t != r && t != s) cs.add(k - m.colptr[c], term(m.cval[k], m.ldout[t], m.ldin[c]));
    }
    return cs.result();
  };
  double inx_r = col_old(r), inx_s = col_old(s);
  // merged row: row r entries (M_rt + M_st), then row s entries absent from row r
  double ldo = clog2((double)(m.dout[r] + m.dout[s]));
  CSum na, nb;
  for (int64_t k = m.rowptr[r]; k < m.rowptr[r + 1]; ++k) {
    int32_t t = m.col[k];
    if (exact && (t == r || t == s)) continue;
    na.add(k - m.rowptr[r], term(m.val[k] + m.row_get(s, t), m.ldin[t], ldo));
  }
  if (s != r)
    for (int64_t k = m.rowptr[s]; k < m.rowptr[s + 1]; ++k) {
      int32_t t = m.col[k];
      if (exact && (t == r || t == s)) continue;
      if (m.row_get(r, t) == 0) nb.add(k - m.rowptr[s], term(m.val[k], m.ldin[t], ldo));
    }
  double new_out = na.result() + nb.result();
  // merged column, t not in {r, s}
  double ldi = clog2((double)(m.din[r] + m.din[s]));
  if (exact) {
    const int64_t X = m.row_get(r, r) + m.row_get(r, s) + m.row_get(s, r) + m.row_get(s, s);
    new_out = new_out + (X > 0 ? term(X, ldi, ldo) : 0.0);
  }
  CSum ia, ib;
  for (int64_t k = m.colptr[r]; k < m.colptr[r + 1]; ++k) {
    int32_t t = m.row[k];
    if (t == r || t == s) continue;
    ia.add(k - m.colptr[r], term(m.cval[k] + m.col_get(s, t), m.ldout[t], ldi));
  }
  if (s != r)
    for (int64_t k = m.colptr[s]; k < m.colptr[s + 1]; ++k) {
      int32_t t = m.row[k];
      if (t == r || t == s) continue;
      if (m.col_get(r, t) == 0) ib.add(k - m.colptr[s], term(m.cval[k], m.ldout[t], ldi));
    }
  double new_in = ia.result() + ib.result();
  double d = OUT[r] + OUT[s];
  d = d + (inx_r + inx_s);
  d = d - new_out;
  d = d - new_in;
  return d;
}

// find_best_warp (kernels.hpp:452-491): 32 lanes, pad with +MAX/0, shuffle-down tree,
// ties go to the other lane.
void argmin_tree(const double *dS, const int32_t *prop, int np, double *best_dS, int32_t *best) {
  double v[32];
  int32_t b[32];
  for (int l = 0; l < 32; ++l) {
    if (l < np) { v[l] = dS[l]; b[l] = prop[l]; }
    else { v[l] = std::numeric_limits<double>::max(); b[l] = 0; }
  }
  for (int off = 16; off >= 1; off >>= 1)
    for (int l = 0; l < off; ++l) {
      double a = v[l], o = v[l + off];
      double mn = std::isnan(a) ? o : (o < a ? o : a);
      if (mn == o) b[l] = b[l + off];
      v[l] = mn;
    }
  *best_dS = v[0];
  *best = b[0];
}

// Total order on doubles (sign-magnitude -> unsigned), as a radix sort on float bits orders.
inline uint64_t order_key(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

int32_t find_root(std::vector<int32_t> &par, int32_t x) {
  int32_t r = x;
  while (par[r] != r) r = par[r];
  while (par[x] != r) { int32_t n = par[x]; par[x] = r; x = n; }
  return r;
}

// carry_out_best_merges (partition.hpp:26-78) as union-find.  Returns merges done.
int32_t carry_out_merges(const std::vector<int32_t> &best, const std::vector<double> &dS_best,
                         int32_t B, int32_t B_to_merge, std::vector<int32_t> &map) {
  std::vector<int32_t> order(B);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return order_key(dS_best[a]) < order_key(dS_best[b]);
  });
  std::vector<int32_t> par(B);
  std::iota(par.begin(), par.end(), 0);
  int32_t merged = 0, c = 0;
  while (merged < B_to_merge && c < B) {
    int32_t from = order[c++];
    int32_t to = find_root(par, best[from]);
    if (to != from) { par[from] = to; ++merged; }
  }
  std::vector<int32_t> rank(B, -1);
  int32_t nb = 0;
  for (int32_t x = 0; x < B; ++x)
    if (find_root(par, x) == x) rank[x] = nb++;
  map.resize(B);
  for (int32_t x = 0; x < B; ++x) map[x] = rank[find_root(par, x)];
  return merged;
}

int32_t merge_phase(const G &g, const orc_options &o, const Rng &rng, int32_t *lab, const BM &m,
                    int32_t B_to_merge, uint32_t phase, int32_t *prop, double *dS, int32_t *best,
                    double *dS_best) {
  int32_t B = m.B;
  int np = o.num_proposals;
  std::vector<double> OUT = row_terms(m);
  for (int32_t r = 0; r < B; ++r)
    for (int i = 0; i < np; ++i) {
      int64_t idx = (int64_t)r * np + i;
      prop[idx] = merge_propose(m, rng, phase, r, np, i);
      dS[idx] = merge_dS(m, OUT, r, prop[idx], o.delta_mode == 1);
    }
  std::vector<int32_t> bst(B);
  std::vector<double> bds(B);
  for (int32_t r = 0; r < B; ++r)
    argmin_tree(dS + (int64_t)r * np, prop + (int64_t)r * np, np, &bds[r], &bst[r]);
  if (best) std::copy(bst.begin(), bst.end(), best);
  if (dS_best) std::copy(bds.begin(), bds.end(), dS_best);
  std::vector<int32_t> map;
  int32_t merged = carry_out_merges(bst, bds, B, B_to_merge, map);
  for (int64_t v = 0; v < g.N; ++v) lab[v] = map[lab[v]];
  return B - merged;
}

// ------------------------------------------------------------------ nodal batch
struct Scratch {
  std::vector<int64_t> a, c;
  std::vector<int64_t> fa, fc;
  std::vector<int64_t> sa, sc;
  int64_t stamp = 0;
  void ensure(int32_t B) {
    if ((int32_t)a.size() < B) {
      a.assign(B, 0); c.assign(B, 0); fa.assign(B, 0); fc.assign(B, 0);
      sa.assign(B, -1); sc.assign(B, -1); stamp = 0;
    }
  }
};

// Delta-S of moving one vertex v from r to s against the batch-start state.
// Formulas: kernels.hpp:494-643 (old terms, accept), 1153-1726 (new rows / columns), with the
// node-only entries run-merged.  Entries of rows r, s and columns r, s that v does not touch
// keep their count and change only through the row / column degree, so their old-minus-new
// difference is (log2 d_old - log2 d_new) * W, W = their summed count; the remaining entries
// ("slots": v's out-blocks, r and s for the rows; v's in-blocks outside {r, s} for the
// columns) are evaluated term by term.  Slot order is the first-occurrence position in v's
// out (in) adjacency, then r, then s.
//
// Exact form (delta_mode 1): the new rows r and s use the new column degrees of r and s, and
// a self loop of weight sig moves with v, from (r, r) to (s, s):
//   R_t = M_rt - a_t - [t=r](c_r - sig) + [t=s](c_r - sig)
//   S_t = M_st + a_t - [t=r](c_s + sig) + [t=s](c_s + sig)
// Every entry of rows r, s and columns r, s then enters with its new degrees, so dS is the
// change of S.
//
// With *log2H set, also the log2 Hastings ratio of Metropolis-Hastings acceptance: the
// nodal sampler proposes s for v with probability
//   p(r -> s) = sum_t (k_t / k_v) (M_ts + M_st + 1) / (d_t + B)
// (neighbour block t with probability k_t / k_v, then a uniform block with probability
// B / (d_t + B), else a block of t's row and column weighted by M; d_t = d_out_t + d_in_t).
// The sum runs over v's adjacency entries in order (out entries, then in entries, each in
// canonical order); the reverse proposal p(s -> r) uses the state after the move (M', d',
// and the self loop's neighbour block s).  log2H = log2 p(s -> r) - log2 p(r -> s); 0 for a
// vertex without edges (both proposals uniform).
double nodal_vertex(const G &g, const BM &m, const int32_t *lab, int64_t v, int32_t r, int32_t s,
                    Scratch &sc, bool exact, double *log2H) {
  if (r == s) return 0.0;
  ++sc.stamp;
  const int64_t st = sc.stamp;
  auto A = [&](int32_t t) -> int64_t { return sc.sa[t] == st ? sc.a[t] : 0; };
  auto C = [&](int32_t t) -> int64_t { return sc.sc[t] == st ? sc.c[t] : 0; };
  const int64_t o0 = g.optr[v], o1 = g.optr[v + 1], i0 = g.iptr[v], i1 = g.iptr[v + 1];
  for (int64_t j = o0; j < o1; ++j) {
    int32_t t = lab[g.oadj[j]];
    if (sc.sa[t] != st) { sc.sa[t] = st; sc.a[t] = 0; sc.fa[t] = j - o0; }
    sc.a[t] += g.ow[j];
  }
  for (int64_t j = i0; j < i1; ++j) {
    int32_t t = lab[g.iadj[j]];
    if (sc.sc[t] != st) { sc.sc[t] = st; sc.c[t] = 0; sc.fc[t] = j - i0; }
    sc.c[t] += g.iw[j];
  }
  const int64_t c_r = C(r), c_s = C(s);
  const int64_t kout = g.kout[v], kin = g.kin[v];
  int64_t sloop = 0;  // weight of v's self loop
  for (int64_t j = o0; j < o1; ++j)
    if (g.oadj[j] == v) sloop += g.ow[j];
  // the exact delta moves the self loop from (r, r) to (s, s); the reference's form leaves
  // it out of R_t and S_t
  const int64_t sig = exact ? sloop : 0;
  const double l_dr = clog2((double)(m.dout[r] - kout)), l_ds = clog2((double)(m.dout[s] + kout));
  const double l_ir = clog2((double)(m.din[r] - kin)), l_is = clog2((double)(m.din[s] + kin));

  // rows r and s: old terms minus new terms of the touched columns t
  //   new row r: R_t = M_rt - a_t - [t=r]c_r + [t=s]c_r        (kernels.hpp:1444-1581)
  //   new row s: S_t = M_st + a_t - [t=r]c_s + [t=s]c_s        (kernels.hpp:1584-1726)
  CSum drow;
  int64_t sum_r = 0, sum_s = 0;
  auto row_slot = [&](int64_t k, int32_t t) {
    const int64_t Mr = m.row_get(r, t), Ms = m.row_get(s, t), a = A(t);
    const double or_ = Mr > 0 ? term(Mr, m.ldout[r], m.ldin[t]) : 0.0;
    const double os_ = Ms > 0 ? term(Ms, m.ldout[s], m.ldin[t]) : 0.0;
    const int64_t R = Mr - a - (t == r ? c_r - sig : 0) + (t == s ? c_r - sig : 0);
    const int64_t S = Ms + a - (t == r ? c_s + sig : 0) + (t == s ? c_s + sig : 0);
    const double ln = !exact ? m.ldin[t] : t == r ? l_ir : t == s ? l_is : m.ldin[t];
    const double nr = R > 0 ? term(R, ln, l_dr) : 0.0;
    const double ns = S > 0 ? term(S, ln, l_ds) : 0.0;
    drow.add(k, ((or_ + os_) - nr) - ns);
    sum_r += Mr;
    sum_s += Ms;
  };
  const int64_t no = o1 - o0, ni = i1 - i0;
  for (int64_t j = 0; j < no; ++j) {
    int32_t t = lab[g.oadj[o0 + j]];
    if (sc.fa[t] == j) row_slot(j, t);
  }
  if (A(r) == 0) row_slot(no, r);
  if (A(s) == 0) row_slot(no + 1, s);
  const int64_t W_r = m.dout[r] - sum_r, W_s = m.dout[s] - sum_s;
  const double Y_r = W_r > 0 ? (m.ldout[r] - l_dr) * (double)W_r : 0.0;
  const double Y_s = W_s > 0 ? (m.ldout[s] - l_ds) * (double)W_s : 0.0;

  // columns r and s, t not in {r, s}
  //   new column r: M_tr - c_t   (kernels.hpp:1153-1250);  new column s: M_ts + c_t (1253-1362)
  CSum dcol;
  int64_t csum_r = 0, csum_s = 0;
  for (int64_t j = 0; j < ni; ++j) {
    int32_t t = lab[g.iadj[i0 + j]];
    if (sc.fc[t] != j || t == r || t == s) continue;
    const int64_t Mr = m.col_get(r, t), Ms = m.col_get(s, t), c = C(t);
    const double ir = Mr > 0 ? term(Mr, m.ldout[t], m.ldin[r]) : 0.0;
    const double is = Ms > 0 ? term(Ms, m.ldout[t], m.ldin[s]) : 0.0;
    const int64_t wr = Mr - c, ws = Ms + c;
    const double nr = wr > 0 ? term(wr, m.ldout[t], l_ir) : 0.0;
    const double ns = ws > 0 ? term(ws, m.ldout[t], l_is) : 0.0;
    dcol.add(j, ((ir + is) - nr) - ns);
    csum_r += Mr;
    csum_s += Ms;
  }
  const int64_t W2_r = m.din[r] - m.col_get(r, r) - m.col_get(r, s) - csum_r;
  const int64_t W2_s = m.din[s] - m.col_get(s, r) - m.col_get(s, s) - csum_s;
  const double Y2_r = W2_r > 0 ? (m.ldin[r] - l_ir) * (double)W2_r : 0.0;
  const double Y2_s = W2_s > 0 ? (m.ldin[s] - l_is) * (double)W2_s : 0.0;

  // calculate_dS_overall_nodal (kernels.hpp:600-643): accept iff dS < 0
  double d = drow.result() - Y_r;
  d = d - Y_s;
  d = d + dcol.result();
  d = d - Y2_r;
  d = d - Y2_s;

  if (log2H) {
    const int64_t Bq = m.B, kv = kout + kin;
    const int64_t Mrs = m.row_get(r, s), Msr = m.row_get(s, r), Mss = m.row_get(s, s);
    // M'_rr, M'_rs, M'_sr of the block matrix after the move (the self loop goes from
    // (r, r) to (s, s) whatever the delta-S form)
    const int64_t Rr = m.row_get(r, r) - A(r) - c_r + sloop, Rs = Mrs - A(s) + c_r - sloop;
    const int64_t Sr = Msr + A(r) - c_s - sloop;
    const int64_t dr = m.dout[r] + m.din[r], ds = m.dout[s] + m.din[s];
    auto hterms = [&](int32_t u, int64_t w, double &fw, double &bk) {
      const int32_t t = lab[u];
      int64_t nf, df, nb, db;
      if (t == r) { nf = Mrs + Msr + 1; df = dr + Bq; }
      else if (t == s) { nf = 2 * Mss + 1; df = ds + Bq; }
      else { nf = m.col_get(s, t) + m.row_get(s, t) + 1; df = m.dout[t] + m.din[t] + Bq; }
      if (u == v || t == s) { nb = Sr + Rs + 1; db = ds + kv + Bq; }      // t' = s
      else if (t == r) { nb = 2 * Rr + 1; db = dr - kv + Bq; }             // t' = r
      else { nb = (m.col_get(r, t) - C(t)) + (m.row_get(r, t) - A(t)) + 1; db = df; }
      fw = ((double)w * (double)nf) / (double)df;
      bk = ((double)w * (double)nb) / (double)db;
    };
    CSum fo, fi, bo, bi;
    for (int64_t j = 0; j < no; ++j) {
      double f, b;
      hterms(g.oadj[o0 + j], g.ow[o0 + j], f, b);
      fo.add(j, f);
      bo.add(j, b);
    }
    for (int64_t j = 0; j < ni; ++j) {
      double f, b;
      hterms(g.iadj[i0 + j], g.iw[i0 + j], f, b);
      fi.add(j, f);
      bi.add(j, b);
    }
    const double pf = fo.result() + fi.result(), pb = bo.result() + bi.result();
    *log2H = no + ni == 0 ? 0.0 : clog2(pb) - clog2(pf);
  }
  return d;
}

// Metropolis-Hastings decision: accept iff u < exp(-beta dS) H, compared in log2 with
// u = (x + 1) 2^-32 from the vertex's third Philox word.
bool mh_accepts(double d, double log2H, double beta, uint32_t x) {
  const double L = ((-beta) * d) * 1.4426950408889634 + log2H;
  return clog2((double)x + 1.0) - 32.0 < L;
}

void nodal_batch(const G &g, const orc_options &o, const Rng &rng, int32_t *lab, const BM &m,
                 int32_t b, uint32_t phase, int32_t *prop, double *dS, double *dE, Scratch &sc,
                 int nthr = 1) {
  const int64_t N = g.N, BS = N / o.num_batches, lo = (int64_t)b * BS;
  const int32_t B = m.B;
  // per-block neighbour sample and coin, shared by all vertices (Q5, partition.hpp:485-523)
  std::vector<int32_t> sof(B);
  std::vector<uint8_t> coin(B);
  for (int32_t x = 0; x < B; ++x) {
    uint32_t y[4];
    rng.draw((uint32_t)x, phase, TAG_BLOCK, y);
    int64_t deg = m.dout[x] + m.din[x];
    sof[x] = deg == 0 ? 0 : walk_block(m, x, mulhi(y[0], deg), -1, 0);
    coin[x] = coin_random(y[1], deg, B) ? 1 : 0;
  }
  sc.ensure(B);
  std::vector<int32_t> ps(BS);
  std::vector<double> ds(BS);
  std::vector<Scratch> extra(nthr > 1 ? nthr - 1 : 0);
  par_ranges(nthr, BS, [&](int t, int64_t i0, int64_t i1) {
  Scratch &tsc = t == 0 ? sc : extra[t - 1];
  tsc.ensure(B);
  for (int64_t i = i0; i < i1; ++i) {
    int64_t v = lo + i;
    uint32_t x[4];
    rng.draw((uint32_t)v, phase, TAG_NODE, x);
    int64_t tot = g.kout[v] + g.kin[v];
    int32_t s;
    if (tot == 0) {
      s = (int32_t)mulhi(x[0], B);
    } else {
      uint64_t target = mulhi(x[1], tot), cum = 0;
      int64_t nbr = -1;
      for (int64_t j = g.optr[v]; j < g.optr[v + 1] && nbr < 0; ++j) {
        cum += (uint64_t)g.ow[j];
        if (cum > target) nbr = g.oadj[j];
      }
      for (int64_t j = g.iptr[v]; j < g.iptr[v + 1] && nbr < 0; ++j) {
        cum += (uint64_t)g.iw[j];
        if (cum > target) nbr = g.iadj[j];
      }
      int32_t u = lab[nbr];
      s = coin[u] ? (int32_t)mulhi(x[0], B) : sof[u];
    }
    const bool mh = o.accept == 1;
    double lh = 0.0;
    double d = nodal_vertex(g, m, lab, v, lab[v], s, tsc, o.delta_mode == 1, mh ? &lh : nullptr);
    if (!mh) {
      ps[i] = s;
      ds[i] = d < 0.0 ? d : 0.0;
    } else {
      const bool acc = s != lab[v] && mh_accepts(d, lh, o.beta, x[2]);
      ps[i] = acc ? s : lab[v];
      ds[i] = acc ? d : 0.0;
    }
  }
  });
  // greedy: accepted moves are those with dS < 0; MH: prop holds every vertex's new label
  for (int64_t i = 0; i < BS; ++i)
    if (o.accept == 1 || ds[i] < 0.0) lab[lo + i] = ps[i];
  if (prop) std::copy(ps.begin(), ps.end(), prop);
  if (dS) std::copy(ds.begin(), ds.end(), dS);
  *dE = csum_array(ds.data(), BS);
}

// compute_overall_entropy_cuda (partition.hpp:665-704) in FP64, canonical order.  model 1
// scales the x ln x term by E (the description length of Peixoto's degree-corrected SBM).
double mdl(const BM &m, int64_t N, int64_t E, int32_t model) {
  std::vector<double> t(m.nnz);
  for (int32_t r = 0; r < m.B; ++r)
    for (int64_t k = m.rowptr[r]; k < m.rowptr[r + 1]; ++k)
      t[k] = term(m.val[k], m.ldout[r], m.ldin[m.col[k]]);
  double data_S = csum_array(t.data(), m.nnz);
  const double LN2 = 0.6931471805599453094;
  double B = (double)m.B, dE = (double)E;
  double x = (B * B) / dE;
  double mt;
  if (model == 1)
    mt = dE * ((1.0 + x) * (clog2(1.0 + x) * LN2) - x * (clog2(x) * LN2)) +
         (double)N * (clog2(B) * LN2);
  else
    mt = dE * (1.0 + x) * (clog2(1.0 + x) * LN2) - x * (clog2(x) * LN2) +
         (double)N * (clog2(B) * LN2);
  return mt - data_S;
}

uint64_t fnv(const void *p, int64_t n, uint64_t h = 1469598103934665603ull) {
  const unsigned char *c = (const unsigned char *)p;
  for (int64_t i = 0; i < n; ++i) { h ^= c[i]; h *= 1099511628211ull; }
  return h;
}


}  // namespace

// =================================================================== C ABI
extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
}

double orc_log2(double x) { return clog2(x); }

void orc_log2_tables(double *inv_c, double *log2_c, double *poly) {
  const Log2Tables &T = tables();
  std::memcpy(inv_c, T.inv, sizeof T.inv);
  std::memcpy(log2_c, T.lg, sizeof T.lg);
  std::memcpy(poly, T.poly, sizeof T.poly);
}

double orc_csum64(const double *x, int64_t n) {
  CSum c;
  for (int64_t i = 0; i < n; ++i) c.add(i, x[i]);
  return c.result();
}

double orc_csum_array(const double *x, int64_t n) { return csum_array(x, n); }

uint64_t orc_fnv1a(const void *p, int64_t nbytes, uint64_t h) { return fnv(p, nbytes, h); }

int64_t orc_blockmodel(int64_t N, int64_t E, const int32_t *out_ptr, const int32_t *out_adj,
                       const int32_t *out_w, const int32_t *labels, int32_t B, int64_t *rowptr,
                       int32_t *col, int64_t *val, int64_t *colptr, int32_t *row, int64_t *cval,
                       int64_t *dout, int64_t *din) {
  if (N < 0 || E < 0 || B <= 0 || out_ptr[N] != E) return -1;
  for (int64_t v = 0; v < N; ++v)
    if (labels[v] < 0 || labels[v] >= B) return -1;
  BM m;
  build_bm(N, out_ptr, out_adj, out_w, labels, B, m);
  std::copy(m.rowptr.begin(), m.rowptr.end(), rowptr);
  std::copy(m.colptr.begin(), m.colptr.end(), colptr);
  std::copy(m.col.begin(), m.col.end(), col);
  std::copy(m.val.begin(), m.val.end(), val);
  std::copy(m.row.begin(), m.row.end(), row);
  std::copy(m.cval.begin(), m.cval.end(), cval);
  std::copy(m.dout.begin(), m.dout.end(), dout);
  std::copy(m.din.begin(), m.din.end(), din);
  return m.nnz;
}

int32_t orc_merge_phase(const orc_graph *gg, const orc_options *o, int32_t *labels, int32_t B,
                        int32_t B_to_merge, uint32_t m_phase, int32_t *prop, double *dS,
                        int32_t *best, double *dS_best) {
  G g(gg);
  Rng rng{(uint32_t)o->seed, (uint32_t)(o->seed >> 32), o->chain};
  BM m;
  build_bm(g.N, g.optr, g.oadj, g.ow, labels, B, m);
  std::vector<int32_t> p((int64_t)B * o->num_proposals);
  std::vector<double> d((int64_t)B * o->num_proposals);
  int32_t nb = merge_phase(g, *o, rng, labels, m, B_to_merge, m_phase, p.data(), d.data(), best,
                           dS_best);
  if (prop) std::copy(p.begin(), p.end(), prop);
  if (dS) std::copy(d.begin(), d.end(), dS);
  return nb;
}

void orc_nodal_batch(const orc_graph *gg, const orc_options *o, int32_t *labels, int32_t B,
                     int32_t b, uint32_t q, int32_t *prop, double *dS, double *dE) {
  G g(gg);
  Rng rng{(uint32_t)o->seed, (uint32_t)(o->seed >> 32), o->chain};
  BM m;
  build_bm(g.N, g.optr, g.oadj, g.ow, labels, B, m);
  Scratch sc;
  nodal_batch(g, *o, rng, labels, m, b, q, prop, dS, dE, sc);
}

void orc_nodal_batch_mt(const orc_graph *gg, const orc_options *o, int32_t *labels, int32_t B,
                        int32_t b, uint32_t q, int32_t *prop, double *dS, double *dE,
                        int32_t threads) {
  G g(gg);
  Rng rng{(uint32_t)o->seed, (uint32_t)(o->seed >> 32), o->chain};
  BM m;
  build_bm(g.N, g.optr, g.oadj, g.ow, labels, B, m, threads);
  Scratch sc;
  nodal_batch(g, *o, rng, labels, m, b, q, prop, dS, dE, sc, threads);
}

double orc_nodal_delta(const orc_graph *gg, const int32_t *labels, int32_t B, int64_t v,
                       int32_t s, int32_t exact, double *log2H) {
  G g(gg);
  BM m;
  build_bm(g.N, g.optr, g.oadj, g.ow, labels, B, m);
  Scratch sc;
  sc.ensure(B);
  if (log2H) *log2H = 0.0;
  return nodal_vertex(g, m, labels, v, labels[v], s, sc, exact != 0, log2H);
}

double orc_merge_delta(const orc_graph *gg, const int32_t *labels, int32_t B, int32_t r,
                       int32_t s, int32_t exact) {
  BM m;
  build_bm(gg->N, gg->out_ptr, gg->out_adj, gg->out_w, labels, B, m);
  return merge_dS(m, row_terms(m), r, s, exact != 0);
}

double orc_mdl(const orc_graph *gg, const int32_t *labels, int32_t B, int32_t model) {
  BM m;
  build_bm(gg->N, gg->out_ptr, gg->out_adj, gg->out_w, labels, B, m);
  return mdl(m, gg->N, gg->E, model);
}

int orc_partition(const orc_graph *gg, const orc_options *o, int32_t *labels_out, orc_report *rep,
                  orc_trace_rec *trace, int64_t trace_cap) {
  return orc_partition_mt(gg, o, labels_out, rep, trace, trace_cap, 1);
}

int orc_partition_mt(const orc_graph *gg, const orc_options *o, int32_t *labels_out,
                     orc_report *rep, orc_trace_rec *trace, int64_t trace_cap, int32_t threads) {
  if (!gg || !o || !labels_out || gg->N <= 0) return -1;
  if (o->num_proposals < 1 || o->num_proposals > 32 || o->num_batches < 1 ||
      o->max_nodal_itr < 1 || o->moving_avg_window < 2)
    return -1;
  if (gg->E == 0) {
    // no edges: the model term divides by E (partition.hpp:696-699); one block, S = 0
    std::fill(labels_out, labels_out + gg->N, 0);
    if (rep) {
      *rep = orc_report{};
      rep->B = 1;
    }
    return 0;
  }
  G g(gg);
  Rng rng{(uint32_t)o->seed, (uint32_t)(o->seed >> 32), o->chain};
  const int64_t N = g.N;
  const int np = o->num_proposals;
  std::vector<int32_t> lab(N);
  std::iota(lab.begin(), lab.end(), 0);
  int64_t ntr = 0;
  auto rec = [&](int32_t kind, int32_t B, int32_t a, int32_t b, uint64_t h1, uint64_t h2, double x,
                 double y) {
    if (trace && ntr < trace_cap) trace[ntr] = orc_trace_rec{kind, B, a, b, h1, h2, x, y};
    ++ntr;
  };

  // partition.hpp:871-875
  int32_t B = (int32_t)N;
  int32_t B_to_merge = (int32_t)((float)B * o->reduction_rate);
  // golden-section slots large, med, small (graph.hpp:57-70: B = 0, S = inf), each holding
  // one of three label snapshots
  orc_golden_slot slots[3];
  orc_golden_init(slots);
  std::vector<int32_t> snap[3];
  BM m;
  build_bm(N, g.optr, g.oadj, g.ow, lab.data(), B, m, threads);
  Scratch sc;
  std::vector<int32_t> prop, best;
  std::vector<double> dS, dS_best;
  uint32_t merge_ctr = 0, batch_ctr = 0;
  int32_t levels = 0, sweeps = 0;
  int64_t batches = 0;
  double S = 0.0;
  bool optimal = false;
  const int64_t BS = N / o->num_batches;
  while (!optimal) {
    // block merge   (partition.hpp:900-904)
    prop.resize((int64_t)B * np); dS.resize((int64_t)B * np);
    best.resize(B); dS_best.resize(B);
    int32_t nB = merge_phase(g, *o, rng, lab.data(), m, B_to_merge, merge_ctr, prop.data(),
                             dS.data(), best.data(), dS_best.data());
    rec(1, B, nB, (int32_t)merge_ctr, fnv(prop.data(), 4 * (int64_t)prop.size()),
        fnv(lab.data(), 4 * N), 0.0, 0.0);
    ++merge_ctr;
    B = nB;
    build_bm(N, g.optr, g.oadj, g.ow, lab.data(), B, m, threads);
    // nodal sweeps   (partition.hpp:906-935)
    std::vector<double> dE(o->max_nodal_itr, 0.0);
    std::vector<int32_t> bp(BS);
    std::vector<double> bd(BS);
    for (int itr = 0; itr < o->max_nodal_itr; ++itr) {
      for (int b = 0; b < o->num_batches; ++b) {
        double e = 0.0;
        std::vector<int32_t> before(lab.begin() + (int64_t)b * BS, lab.begin() + (int64_t)(b + 1) * BS);
        nodal_batch(g, *o, rng, lab.data(), m, b, batch_ctr, bp.data(), bd.data(), &e, sc, threads);
        build_bm(N, g.optr, g.oadj, g.ow, lab.data(), B, m, threads);
        int64_t moved = 0;  // vertices of the batch whose label changed
        for (int64_t i = 0; i < BS; ++i) moved += lab[(int64_t)b * BS + i] != before[i];
        rec(2, B, itr, b, fnv(bp.data(), 4 * BS), fnv(lab.data(), 4 * N), e, (double)moved);
        ++batch_ctr; ++batches;
        dE[itr] = dE[itr] + e;
      }
      S = mdl(m, N, g.E, o->model);
      ++sweeps;
      std::vector<int32_t> v32(m.val.begin(), m.val.end());
      rec(3, B, itr, 0, fnv(m.col.data(), 4 * m.nnz), fnv(v32.data(), 4 * m.nnz), S, dE[itr]);
      const int bracketed = std::isfinite(slots[0].S) && std::isfinite(slots[1].S) &&
                            std::isfinite(slots[2].S);
      if (orc_converged(dE.data(), itr, o->moving_avg_window, S, bracketed, o->threshold1,
                        o->threshold2))
        break;
    }
    // golden section   (partition.hpp:80-322); the labels of the level just finished go
    // to snapshot `store`, the next level starts from snapshot `restore`
    const orc_golden_slot P{B, B_to_merge, S, -1, 0};
    orc_golden_slot next;
    int32_t store = 0, restore = 0, opt = 0;
    orc_golden_step(slots, &P, o->reduction_rate, &store, &restore, &next, &opt);
    snap[store] = lab;
    optimal = opt != 0;
    B = next.B;
    B_to_merge = next.B_to_merge;
    S = next.S;
    lab = snap[restore];
    build_bm(N, g.optr, g.oadj, g.ow, lab.data(), B, m, threads);
    ++levels;
    rec(4, B, B_to_merge, optimal ? 1 : 0, fnv(lab.data(), 4 * N), 0, S, 0.0);
    if (o->max_levels > 0 && levels >= o->max_levels) break;
  }
  std::copy(lab.begin(), lab.end(), labels_out);
  if (rep) {
    rep->B = B; rep->levels = levels; rep->merge_phases = (int32_t)merge_ctr; rep->sweeps = sweeps;
    rep->batches = batches; rep->S = S; rep->n_trace = ntr;
  }
  return 0;
}

void orc_golden_init(orc_golden_slot slots[3]) {
  // OldData() (graph.hpp:62-69): B = 0 and S = inf in every slot; slot k owns snapshot k
  for (int32_t k = 0; k < 3; ++k)
    slots[k] = orc_golden_slot{0, 0, std::numeric_limits<double>::infinity(), k, 0};
}

int32_t orc_converged(const double *dE, int32_t itr, int32_t W, double S, int32_t bracketed,
                      double threshold1, double threshold2) {
  // partition.hpp:918-934: from itr = W-1 on, the mean of the W-1 sweeps before itr (the
  // current sweep is not in the window) against threshold1 * S until all three slots hold a
  // finite S, threshold2 * S after
  if (itr < W - 1) return 0;
  double mean = 0.0;
  for (int32_t i = itr - W + 1; i < itr; ++i) mean += dE[i];
  mean /= (double)(W - 1);
  return (-mean < (bracketed ? threshold2 : threshold1) * S) ? 1 : 0;
}

void orc_golden_step(orc_golden_slot slots[3], const orc_golden_slot *P, float rate,
                     int32_t *store, int32_t *restore, orc_golden_slot *next, int32_t *optimal) {
  // partition.hpp:89-161: the level's result P replaces med (med moving to large or small,
  // their snapshot buffers exchanged) when it is no worse, else it becomes small or large
  orc_golden_slot &large = slots[0], &med = slots[1], &small = slots[2];
  orc_golden_slot *dst;
  if (P->S <= med.S) {
    if (med.B > P->B) std::swap(large, med);
    else std::swap(small, med);
    dst = &med;
  } else {
    dst = (med.B > P->B) ? &small : &large;
  }
  dst->B = P->B;
  dst->B_to_merge = P->B_to_merge;
  dst->S = P->S;
  *store = dst->snap;
  // partition.hpp:163-320: restore a slot and choose the next B_to_merge
  const orc_golden_slot *from;
  int32_t Bm, opt = 0;
  if (std::isinf(small.S)) {
    from = &med;
    Bm = (int32_t)((float)med.B * rate);  // (int)P.B * rate: a float product, truncated
    if (Bm == 0) opt = 1;
  } else if (large.B - small.B == 2) {
    opt = 1;
    from = &med;
    Bm = med.B_to_merge;
  } else if ((large.B - med.B) >= (med.B - small.B)) {
    const int32_t next_B = med.B + (int32_t)std::round((double)(large.B - med.B) * 0.618);
    from = &large;
    Bm = large.B - next_B;
  } else {
    const int32_t next_B = small.B + (int32_t)std::round((double)(med.B - small.B) * 0.618);
    from = &med;
    Bm = med.B - next_B;
  }
  // not in the reference: a bracket step of no merges (two slots with the same B, possible
  // when a merge phase runs out of candidates) ends the search instead of repeating the level
  if (Bm <= 0 && !opt) opt = 1;
  *restore = from->snap;
  *next = orc_golden_slot{from->B, Bm, from->S, from->snap, 0};
  *optimal = opt;
}

int32_t orc_carry_out_merges(const int32_t *best, const double *dS_best, int32_t B,
                             int32_t B_to_merge, int32_t *map_out) {
  std::vector<int32_t> b(best, best + B), map;
  std::vector<double> d(dS_best, dS_best + B);
  int32_t merged = carry_out_merges(b, d, B, B_to_merge, map);
  std::copy(map.begin(), map.end(), map_out);
  return merged;
}

void orc_argmin_tree(const double *dS, const int32_t *prop, int32_t np, double *best_dS,
                     int32_t *best) {
  argmin_tree(dS, prop, np, best_dS, best);
}

void orc_pairwise(int64_t N, const int32_t *truth, const int32_t *given, double *precision,
                  double *recall) {
  // evaluate.hpp:130-181: sum_ij C(n_ij,2) over sum_i C(a_i,2) (recall), sum_j C(b_j,2) (precision)
  int32_t Bt = 0, Bg = 0;
  for (int64_t i = 0; i < N; ++i) {
    if (truth[i] + 1 > Bt) Bt = truth[i] + 1;
    if (given[i] + 1 > Bg) Bg = given[i] + 1;
  }
  std::vector<int64_t> a(Bt, 0), b(Bg, 0);
  std::vector<std::pair<int64_t, int64_t>> pairs;
  pairs.reserve(N);
  for (int64_t i = 0; i < N; ++i) {
    if (truth[i] < 0) continue;
    a[truth[i]]++; b[given[i]]++;
    pairs.emplace_back(truth[i], given[i]);
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
}

}  // extern "C"
