// Host orchestration of the SBP hot path on one MI355X and the C ABI of gsap.h.
//
// Control flow follows partition() of GSAP/partition.hpp:852-954:
//   level loop: block merge (propose_block_merge + carry_out_best_merges, :900-904)
//               -> up to max_nodal_itr sweeps of num_batches nodal batches, each followed
//                  by a block-matrix rebuild (:912-916)
//               -> total description length + convergence test (:918-934)
//               -> golden-section bracket update over B (:938, :80-322)
// One host thread, one HIP stream, one host<->device sync per sweep (the convergence test
// needs S) and one per merge phase (the greedy merge order is a sequential union-find on
// the host, like the reference's host loop).  The host learns that a sweep or a merge phase
// has finished by polling a sequence number that a one-thread kernel stores into coherent
// host memory after the results (no interrupt-driven wake-up per sweep).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gsap/gsap.h"
#include "engine.h"

using namespace gsap;

namespace {

struct Err : std::runtime_error {
  int code;
  Err(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPX(call)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw Err(GSAP_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

uint64_t fnv(const void *p, int64_t n, uint64_t h = 1469598103934665603ull) {
  const unsigned char *c = (const unsigned char *)p;
  for (int64_t i = 0; i < n; ++i) { h ^= c[i]; h *= 1099511628211ull; }
  return h;
}

// Device allocation that remembers its capacity.
struct DBuf {
  void *p = nullptr;
  size_t cap = 0;
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (bytes == 0) return;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();  // the context stays usable: no stale error for the next call
      throw Err(GSAP_ERR_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    cap = bytes;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T *as() const { return (T *)p; }
};

// Pinned host allocation that remembers its capacity.
struct HPin {
  void *p = nullptr;
  size_t cap = 0;
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    release();
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw Err(GSAP_ERR_MEMORY, "hipHostMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    cap = bytes;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T *as() const { return (T *)p; }
};

// Host memory the device writes with k_publish: results, the block matrix's entry count and
// the sequence number of the last publication.
struct HostPub {
  double v[64];
  int32_t nnz;
  uint32_t seq;
};

}  // namespace

// Optional per-kernel-class device timing with HIP events on the work stream.
struct Prof {
  bool on = false;
  struct Span { int k; hipEvent_t a, b; };
  std::vector<hipEvent_t> pool;
  std::vector<Span> pend;
  double ms[8] = {};
  int64_t calls[8] = {};
  hipEvent_t ev() {
    if (pool.empty()) {
      hipEvent_t e;
      HIPX(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  uint32_t mask = 0xFFu;
  int begin(int k, hipStream_t st) {
    if (!on || !((mask >> k) & 1u)) return -1;
    Span s{k, ev(), nullptr};
    HIPX(hipEventRecord(s.a, st));
    pend.push_back(s);
    return (int)pend.size() - 1;
  }
  void end(int i, hipStream_t st) {
    if (i < 0) return;
    pend[i].b = ev();
    HIPX(hipEventRecord(pend[i].b, st));
  }
  // after a stream synchronisation
  void flush() {
    for (Span &s : pend) {
      float t = 0;
      if (s.b) {
        HIPX(hipEventElapsedTime(&t, s.a, s.b));
        ms[s.k] += t;
        calls[s.k]++;
        pool.push_back(s.b);
      }
      pool.push_back(s.a);
    }
    pend.clear();
  }
  void reset(bool enable) {
    flush();
    on = enable;
    for (int k = 0; k < 8; ++k) { ms[k] = 0; calls[k] = 0; }
  }
  void destroy() {
    flush();
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
  }
};

// A level's labelling at the end of its sweeps and the batches the level ran (autotuner)
struct LevelSnap {
  int32_t B = 0;
  int64_t batches = 0;
  std::vector<int32_t> lab;
};

struct gsap_ctx {
  Prof prof;
  std::vector<LevelSnap> *level_rec = nullptr;  // when set, partition_impl appends snapshots
  int device = 0;
  hipStream_t st = nullptr;
  std::string err;
  std::vector<gsap_trace_rec> trace;
  // graph
  DBuf optr, oadj, ow, iptr, iadj, iw, kout, kin, esrc, idst, sloop, hoff, hbo, hbi, hpool;
  // labels and golden-section snapshots
  DBuf lab, snap[3];
  // block matrix
  DBuf rowptr, colptr, col, val, erow, row, cval, dout, din, ldout, ldin, dense, gcol, grow;
  // the alternate CSR/CSC set of the sparse update after a nodal batch, and its long-segment
  // runs
  DBuf rowptr2, colptr2, col2, val2, erow2, row2, cval2, gcol2, grow2, rk, rw, mwork;
  // rebuild scratch
  DBuf k0, k1, tmp, partials, soff, su, suoff, lists, ncl, cntc, vrel, chunk_seg, lab16, colsum;
  // sweep / merge
  DBuf OUT, sof, coin, prop, dS, best, dSbest, map, tiles, res;
  // device batch counter (Philox phase of the sweep's first batch); merge-order sort
  DBuf ctr, mk0, mk1, mv0, mv1;
  // mixed form: per-vertex |dS32 - dS64| and evaluation class; batch statistics
  DBuf errv, evcls, mixstat, evstat;
  // the current batch's labels before it was applied (incremental dense rebuild)
  DBuf olab;
  // the nodal proposal pass's class lists and their counters (NodalPre)
  DBuf dlist, dcnt;
  // per batch, the batch's vertices split by neighbour-count class (nodal_class)
  DBuf vlist;
  std::vector<int64_t> list_off;  // [batch * 4 + class] -> offset into vlist; last = total
  std::vector<uint32_t> big_lds;  // per batch: LDS bytes of the largest class's launch
  // per batch: 4 x (out- plus in-adjacency of its vertices), the most delta entries its
  // sparse block-matrix update can produce
  std::vector<int64_t> delta_bound;
  // set while a sweep is captured as a HIP graph
  bool capturing = false;
  int32_t lists_nb = 0;
  NodalBounds lists_bounds;
  GraphDev g;
  BMBuf bm;
  RebuildScratch rs;
  int64_t N = 0, E = 0;
  int32_t dense_max_B = 24576;
  int32_t tile_max_B = 2048;
  // sub-contexts of batch / chain calls (one per worker thread)
  std::vector<gsap_ctx *> workers;
  // coherent host memory mapped for the device (host and device addresses) and the last
  // sequence number published into it
  HostPub *pub = nullptr, *pub_dev = nullptr;
  uint32_t pub_seq = 0;
  // pinned staging of the merge phase's arrays (delta-S, best partner, order, merge forest)
  HPin stage;
  // the uploaded graph's row pointers on the host (the batch lists are built from them) and
  // whether a graph is resident for gsap_partition_resident_i32
  std::vector<int32_t> hptr_out, hptr_in;
  bool resident = false;
};

namespace {

void set_device(gsap_ctx *c) { HIPX(hipSetDevice(c->device)); }

// --------------------------------------------------------------- graph upload
// f(lo, hi, t) over [0, n) in contiguous chunks on up to 16 host threads (t = chunk index)
template <class F>
void parallel_chunks(int64_t n, int nt, F &&f) {
  if (nt <= 1 || n < 65536) {
    f((int64_t)0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    th.emplace_back([&f, lo, hi, t] { f(lo, hi, t); });
  }
  for (auto &x : th) x.join();
}

int host_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(16u, h));
}

uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

void validate(const gsap_graph_i32 *g) {
  if (!g) throw Err(GSAP_ERR_INVALID, "graph is NULL");
  if (g->N <= 0 || g->E < 0) throw Err(GSAP_ERR_INVALID, "N must be > 0 and E >= 0");
  if (g->N >= (int64_t)1 << 31 || g->E >= (int64_t)1 << 31)
    throw Err(GSAP_ERR_INVALID, "N and E must fit int32");
  if (!g->out_ptr || !g->in_ptr || (g->E > 0 && (!g->out_adj || !g->out_w || !g->in_adj || !g->in_w)))
    throw Err(GSAP_ERR_INVALID, "graph arrays must not be NULL");
  if (g->out_ptr[0] != 0 || g->in_ptr[0] != 0 || g->out_ptr[g->N] != g->E || g->in_ptr[g->N] != g->E)
    throw Err(GSAP_ERR_INVALID, "ptr arrays must start at 0 and end at E");
  const int nt = host_threads();
  std::vector<int> bad(nt, 0);
  parallel_chunks(g->N, nt, [&](int64_t lo, int64_t hi, int t) {
    for (int64_t v = lo; v < hi; ++v)
      if (g->out_ptr[v + 1] < g->out_ptr[v] || g->in_ptr[v + 1] < g->in_ptr[v]) bad[t] = 1;
  });
  for (int x : bad)
    if (x) throw Err(GSAP_ERR_INVALID, "ptr arrays must be non-decreasing");
  // per vertex range: ids in range, positive weights, total weight, and order-independent
  // hash sums of the (source, target, weight) triples of both CSR views (the row and column
  // segments of the block matrix are built from one and the other, so they must agree)
  std::vector<int64_t> wsum(nt, 0);
  std::vector<uint64_t> ho(nt, 0), hi_(nt, 0);
  parallel_chunks(g->N, nt, [&](int64_t lo, int64_t hi, int t) {
    int64_t ws = 0;
    uint64_t a = 0, b = 0;
    int err = 0;
    for (int64_t v = lo; v < hi && !err; ++v) {
      for (int32_t j = g->out_ptr[v]; j < g->out_ptr[v + 1]; ++j) {
        const int32_t u = g->out_adj[j], w = g->out_w[j];
        if (u < 0 || u >= g->N) { err = 1; break; }
        if (w <= 0) { err = 2; break; }
        ws += w;
        a += mix64(((uint64_t)v << 32 | (uint32_t)u) ^ mix64((uint64_t)w));
      }
      for (int32_t j = g->in_ptr[v]; j < g->in_ptr[v + 1] && !err; ++j) {
        const int32_t u = g->in_adj[j], w = g->in_w[j];
        if (u < 0 || u >= g->N) { err = 1; break; }
        if (w <= 0) { err = 2; break; }
        b += mix64(((uint64_t)(uint32_t)u << 32 | (uint64_t)v) ^ mix64((uint64_t)w));
      }
    }
    bad[t] = err;
    wsum[t] = ws;
    ho[t] = a;
    hi_[t] = b;
  });
  for (int x : bad) {
    if (x == 1) throw Err(GSAP_ERR_INVALID, "adjacency id out of range");
    if (x == 2) throw Err(GSAP_ERR_INVALID, "weights must be > 0");
  }
  int64_t W = 0;
  uint64_t HO = 0, HI = 0;
  for (int t = 0; t < nt; ++t) { W += wsum[t]; HO += ho[t]; HI += hi_[t]; }
  if (2 * W >= ((int64_t)1 << 31) - 1)
    throw Err(GSAP_ERR_INVALID, "total edge weight must stay below 2^30");
  if (HO != HI) throw Err(GSAP_ERR_INVALID, "in-CSR and out-CSR hold different edges");
}

template <class T>
void h2d(DBuf &b, const T *src, int64_t n, hipStream_t st) {
  b.ensure(sizeof(T) * (size_t)std::max<int64_t>(n, 1));
  if (n > 0) HIPX(hipMemcpyAsync(b.p, src, sizeof(T) * n, hipMemcpyHostToDevice, st));
}

void upload_graph(gsap_ctx *c, const gsap_graph_i32 *g) {
  validate(g);
  c->resident = false;
  const int64_t N = g->N, E = g->E;
  c->N = N; c->E = E;
  c->hptr_out.assign(g->out_ptr, g->out_ptr + N + 1);
  c->hptr_in.assign(g->in_ptr, g->in_ptr + N + 1);
  hipStream_t st = c->st;
  std::vector<int32_t> kout(N, 0), kin(N, 0), sloop(N, 0);
  const int nt = host_threads();
  std::vector<int32_t> wmx(nt, 1);
  parallel_chunks(N, nt, [&](int64_t lo, int64_t hi, int t) {
    int32_t m = 1;
    for (int64_t v = lo; v < hi; ++v) {
      for (int32_t j = g->out_ptr[v]; j < g->out_ptr[v + 1]; ++j) {
        kout[v] += g->out_w[j];
        m = std::max(m, g->out_w[j]);
        if (g->out_adj[j] == v) sloop[v] += g->out_w[j];
      }
      for (int32_t j = g->in_ptr[v]; j < g->in_ptr[v + 1]; ++j) kin[v] += g->in_w[j];
    }
    wmx[t] = m;
  });
  // vertices of the largest nodal class get tables and slot arrays in the global pool
  // (used when they do not fit the launch's LDS); offsets in int32 words, even
  std::vector<int64_t> hoff(N, -1);
  std::vector<uint8_t> hbo(N, 0), hbi(N, 0);
  int64_t pool = 0;
  // (table sizes for every vertex: options may send any vertex to the workgroup form; those
  // above the one-wave form's limit of 128 neighbours may not fit LDS and get pool space)
  for (int64_t v = 0; v < N; ++v) {
    int64_t no = g->out_ptr[v + 1] - g->out_ptr[v], ni = g->in_ptr[v + 1] - g->in_ptr[v];
    hbo[v] = (uint8_t)big_table_bits(no);
    hbi[v] = (uint8_t)big_table_bits(ni);
    if (nodal_class(std::max(no, ni)) == NCLS - 1) {
      hoff[v] = pool;
      pool += (big_words(no, ni) + 1) & ~(int64_t)1;
    }
  }
  h2d(c->optr, g->out_ptr, N + 1, st);
  h2d(c->iptr, g->in_ptr, N + 1, st);
  h2d(c->oadj, g->out_adj, E, st);
  h2d(c->ow, g->out_w, E, st);
  h2d(c->iadj, g->in_adj, E, st);
  h2d(c->iw, g->in_w, E, st);
  h2d(c->kout, kout.data(), N, st);
  h2d(c->kin, kin.data(), N, st);
  // the owner of every CSR entry, filled on the device from the pointer arrays
  c->esrc.ensure(4 * (size_t)std::max<int64_t>(E, 1));
  c->idst.ensure(4 * (size_t)std::max<int64_t>(E, 1));
  launch_owners(c->optr.as<int32_t>(), N, c->esrc.as<int32_t>(), st);
  launch_owners(c->iptr.as<int32_t>(), N, c->idst.as<int32_t>(), st);
  h2d(c->sloop, sloop.data(), N, st);
  h2d(c->hoff, hoff.data(), N, st);
  h2d(c->hbo, hbo.data(), N, st);
  h2d(c->hbi, hbi.data(), N, st);
  c->hpool.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(pool, 1));
  HIPX(hipStreamSynchronize(st));  // host vectors go out of scope
  GraphDev &d = c->g;
  d.N = N; d.E = E;
  d.optr = c->optr.as<int32_t>(); d.oadj = c->oadj.as<int32_t>(); d.ow = c->ow.as<int32_t>();
  d.iptr = c->iptr.as<int32_t>(); d.iadj = c->iadj.as<int32_t>(); d.iw = c->iw.as<int32_t>();
  d.kout = c->kout.as<int32_t>(); d.kin = c->kin.as<int32_t>(); d.esrc = c->esrc.as<int32_t>();
  d.idst = c->idst.as<int32_t>();
  d.sloop = c->sloop.as<int32_t>();
  d.wmax = *std::max_element(wmx.begin(), wmx.end());
  d.hoff = c->hoff.as<int64_t>(); d.hbits_o = c->hbo.as<uint8_t>(); d.hbits_i = c->hbi.as<uint8_t>();
  d.hpool = c->hpool.as<int32_t>();
  c->resident = true;
}

// The resident graph's row pointers as a host graph view (what build_batch_lists reads)
gsap_graph_i32 resident_ptrs(const gsap_ctx *c) {
  gsap_graph_i32 h{};
  h.N = c->N;
  h.E = c->E;
  h.out_ptr = c->hptr_out.data();
  h.in_ptr = c->hptr_in.data();
  return h;
}

// Vertex lists of every nodal batch by class of max(out, in) neighbour count: <= 16, <= 32,
// <= 128, larger.  The last class gets LDS tables sized for its largest vertex (at most 2048
// slots; beyond that the kernel uses the global pool).
void build_batch_lists(gsap_ctx *c, const gsap_graph_i32 *g, int32_t nb,
                       const NodalBounds &bounds = NodalBounds()) {
  const int64_t N = g->N, BS = N / nb;
  // per batch (one host thread each): its vertices by class, classes in order
  std::vector<std::vector<int32_t>> per(nb);
  std::vector<std::array<int64_t, NCLS>> cnt(nb);
  c->big_lds.assign((size_t)nb, 0);
  parallel_chunks(nb, std::min(nb, host_threads()), [&](int64_t lo, int64_t hi, int) {
    for (int64_t b = lo; b < hi; ++b) {
      std::vector<int32_t> cls((size_t)BS);
      std::array<int64_t, NCLS> n{};
      uint32_t lds = 0;
      for (int64_t i = 0; i < BS; ++i) {
        const int64_t v = b * BS + i;
        const int64_t no = g->out_ptr[v + 1] - g->out_ptr[v];
        const int64_t ni = g->in_ptr[v + 1] - g->in_ptr[v];
        const int k = nodal_class(std::max(no, ni), bounds);
        cls[(size_t)i] = k;
        ++n[k];
        if (k == NCLS - 1) {
          // LDS of the launch: enough for its largest vertex, within the cap
          const int64_t need = 4 * big_words(no, ni);
          if (need <= BIG_LDS_CAP) lds = std::max<uint32_t>(lds, (uint32_t)need);
        }
      }
      std::array<int64_t, NCLS> pos{};
      for (int k = 1; k < NCLS; ++k) pos[k] = pos[k - 1] + n[k - 1];
      per[b].resize((size_t)BS);
      for (int64_t i = 0; i < BS; ++i) per[b][(size_t)pos[cls[(size_t)i]]++] = (int32_t)(b * BS + i);
      cnt[b] = n;
      c->big_lds[b] = lds;
    }
  });
  std::vector<int32_t> lst;
  lst.reserve((size_t)(BS * nb));
  c->list_off.assign((size_t)nb * NCLS + 1, 0);
  for (int32_t b = 0; b < nb; ++b) {
    int64_t o = (int64_t)lst.size();
    for (int k = 0; k < NCLS; ++k) {
      c->list_off[(size_t)b * NCLS + k] = o;
      o += cnt[b][k];
    }
    lst.insert(lst.end(), per[b].begin(), per[b].end());
  }
  c->list_off[(size_t)nb * NCLS] = (int64_t)lst.size();
  c->delta_bound.assign((size_t)nb, 0);
  for (int32_t b = 0; b < nb; ++b) {
    const int64_t lo = b * BS, hi = lo + BS;
    c->delta_bound[b] = 4 * ((int64_t)(g->out_ptr[hi] - g->out_ptr[lo]) +
                             (int64_t)(g->in_ptr[hi] - g->in_ptr[lo]));
  }
  c->vlist.ensure(4 * std::max<size_t>(lst.size(), 1));
  if (!lst.empty())
    HIPX(hipMemcpyAsync(c->vlist.p, lst.data(), 4 * lst.size(), hipMemcpyHostToDevice, c->st));
  HIPX(hipStreamSynchronize(c->st));
  c->lists_nb = nb;
  c->lists_bounds = bounds;
}

void alloc_work(gsap_ctx *c, int32_t np, int32_t dense_max_B) {
  const int64_t N = c->N, E = std::max<int64_t>(c->E, 1);
  c->dense_max_B = dense_max_B;
  if (c->bm.flipped) {
    // the last sparse update left the current matrix in the alternate buffers: they become
    // the primary ones again (the dense buffer's entries are described by them)
    std::swap(c->rowptr, c->rowptr2); std::swap(c->colptr, c->colptr2);
    std::swap(c->col, c->col2); std::swap(c->val, c->val2); std::swap(c->erow, c->erow2);
    std::swap(c->row, c->row2); std::swap(c->cval, c->cval2);
    std::swap(c->gcol, c->gcol2); std::swap(c->grow, c->grow2);
    c->bm.flipped = false;
  }
  c->lab.ensure(4 * N);
  for (auto &s : c->snap) s.ensure(4 * N);
  // the dense buffer's zero-except-current-entries state survives only if neither it nor
  // the arrays that describe its entries move
  const void *before[4] = {c->dense.p, c->erow.p, c->col.p, c->rowptr.p};
  c->rowptr.ensure(4 * (N + 1)); c->colptr.ensure(4 * (N + 1));
  c->col.ensure(4 * E); c->val.ensure(4 * E); c->erow.ensure(4 * E);
  c->row.ensure(4 * E); c->cval.ensure(4 * E);
  c->gcol.ensure(4 * (E / 16 + 2)); c->grow.ensure(4 * (E / 16 + 2));
  c->rowptr2.ensure(4 * (N + 1)); c->colptr2.ensure(4 * (N + 1));
  c->col2.ensure(4 * E); c->val2.ensure(4 * E); c->erow2.ensure(4 * E);
  c->row2.ensure(4 * E); c->cval2.ensure(4 * E);
  c->gcol2.ensure(4 * (E / 16 + 2)); c->grow2.ensure(4 * (E / 16 + 2));
  // delta entries of the sparse update: at most the largest batch's bound (4 x its out- plus
  // in-adjacency, about 2E with 4 batches); the segment arrays k0 / k1 hold 2E words already
  int64_t dmax = 2 * E;
  for (int64_t x : c->delta_bound) dmax = std::max(dmax, x);
  c->rk.ensure(4 * dmax); c->rw.ensure(4 * dmax);
  c->mwork.ensure(4 * (size_t)(N + 8));
  c->dout.ensure(4 * N); c->din.ensure(4 * N);
  c->ldout.ensure(8 * N); c->ldin.ensure(8 * N);
  const int64_t dB = std::min<int64_t>(dense_max_B, N);
  c->dense.ensure(4 * (size_t)std::max<int64_t>(dB * dB, 1));
  const void *after[4] = {c->dense.p, c->erow.p, c->col.p, c->rowptr.p};
  if (!std::equal(before, before + 4, after)) {
    HIPX(hipMemsetAsync(c->dense.p, 0, c->dense.cap, c->st));
    c->bm.dense_B = 0;
  }
  c->k0.ensure(4 * dmax); c->k1.ensure(4 * dmax);
  c->tmp.ensure(4 * std::max(scan_temp_words(2 * N + 1), radix_temp_words(N)));
  // partial planes of the small-B histogram (up to 1024 planes of at most 128 x 128
  // counters) and of the row-tile histogram (up to 8M counters), then the transpose
  const int64_t sb = std::min<int64_t>(dB, 128);
  // (and the transpose of the dense matrix, which the row-tile and incremental paths
  // build up to 8192 blocks: at most 8192 x 8192 words)
  const int64_t tB = std::min<int64_t>(dB, 8192);
  const int64_t pc = std::max<int64_t>(
      std::max<int64_t>(1024 * sb * sb, tB * tB),
      dB > 128 ? std::min<int64_t>(E / 2, (int64_t)32 << 20) : 0);
  c->partials.ensure(4 * pc);
  // column chunk sums of the dense matrix above 8192 blocks
  const int64_t csw = dB > 8192 ? dB * ((dB + 63) / 64) : 1;
  c->colsum.ensure(4 * csw);
  c->soff.ensure(4 * (2 * N + 1));
  c->chunk_seg.ensure(4 * (2 * E / 64 + 2));
  c->su.ensure(4 * 2 * N); c->suoff.ensure(4 * (2 * N + 1)); c->vrel.ensure(4 * 2 * N);
  c->lab16.ensure(2 * N);
  c->lists.ensure(4 * (size_t)SEG_NCL * 2 * N); c->ncl.ensure(4 * 16);
  c->cntc.ensure(4 * (size_t)std::max<int64_t>(seg_cntc_words(N, E), 1));
  c->OUT.ensure(8 * N); c->sof.ensure(4 * N); c->coin.ensure(N);
  c->prop.ensure(4 * (size_t)np * N); c->dS.ensure(8 * (size_t)np * N);
  c->best.ensure(4 * N); c->dSbest.ensure(8 * N); c->map.ensure(4 * N);
  // tile sums: the MDL's (E / 1024 + 1), then one region of BS / 1024 + 1 per batch
  c->tiles.ensure(8 * (E / 1024 + N / 1024 + 2 * 64 + 8));
  c->res.ensure(8 * 64);
  c->ctr.ensure(8);
  c->mk0.ensure(8 * N); c->mk1.ensure(8 * N); c->mv0.ensure(4 * N); c->mv1.ensure(4 * N);
  c->errv.ensure(4 * N); c->evcls.ensure(N); c->mixstat.ensure(32);
  c->evstat.ensure(16 * EV_SLOTS);
  c->olab.ensure(4 * N);
  c->dlist.ensure(16 * (size_t)NCLS * N);
  c->dcnt.ensure(4 * NCLS);
  BMBuf &b = c->bm;
  b.rowptr = c->rowptr.as<int32_t>(); b.colptr = c->colptr.as<int32_t>();
  b.col = c->col.as<int32_t>(); b.val = c->val.as<int32_t>(); b.erow = c->erow.as<int32_t>();
  b.row = c->row.as<int32_t>(); b.cval = c->cval.as<int32_t>();
  b.dout = c->dout.as<int32_t>(); b.din = c->din.as<int32_t>();
  b.ldout = c->ldout.as<double>(); b.ldin = c->ldin.as<double>();
  b.dense = c->dense.as<int32_t>(); b.dense_cap = (int64_t)(c->dense.cap / 4);
  b.gcol = c->gcol.as<int32_t>(); b.grow = c->grow.as<int32_t>();
  b.a_rowptr = c->rowptr2.as<int32_t>(); b.a_colptr = c->colptr2.as<int32_t>();
  b.a_col = c->col2.as<int32_t>(); b.a_val = c->val2.as<int32_t>();
  b.a_erow = c->erow2.as<int32_t>(); b.a_row = c->row2.as<int32_t>();
  b.a_cval = c->cval2.as<int32_t>(); b.a_gcol = c->gcol2.as<int32_t>();
  b.a_grow = c->grow2.as<int32_t>();
  RebuildScratch &r = c->rs;
  r.k0 = c->k0.p; r.k1 = c->k1.p; r.tmp = c->tmp.as<uint32_t>();
  r.tmp_words = (int64_t)(c->tmp.cap / 4);
  r.partials = c->partials.as<int32_t>(); r.partials_cap = pc;
  r.colsum = c->colsum.as<int32_t>(); r.colsum_cap = (int64_t)(c->colsum.cap / 4);
  r.soff = c->soff.as<uint32_t>();
  r.su = c->su.as<uint32_t>(); r.suoff = c->suoff.as<uint32_t>(); r.vrel = c->vrel.as<uint32_t>(); r.chunk_seg = c->chunk_seg.as<uint32_t>(); r.lab16 = c->lab16.as<uint16_t>();
  r.lists = c->lists.as<uint32_t>(); r.lcap = 2 * N; r.ncl = c->ncl.as<uint32_t>();
  r.cntc = c->cntc.as<uint32_t>();
  r.kcap = (int64_t)(c->k0.cap / 4);
  r.rk = c->rk.as<uint32_t>(); r.rw = c->rw.as<uint32_t>(); r.rcap = (int64_t)(c->rk.cap / 4);
  r.mwork = c->mwork.as<uint32_t>();
  r.mwork_list = N;
}

constexpr int32_t SMALL_B_TUNE = 128;  // the small rebuild path's limit (never tuned)
int32_t resolve_dense(const gsap_options &o) { return o.dense_max_B > 0 ? o.dense_max_B : 24576; }
NodalBounds resolve_nodal(const gsap_options &o) {
  NodalBounds b;
  if (o.nodal_max16) b.max16 = o.nodal_max16;
  if (o.nodal_max32) b.max32 = o.nodal_max32;
  if (o.nodal_maxwave) b.maxwave = o.nodal_maxwave;
  return b;
}
// the segment-class choices of the options into the rebuild scratch
void apply_seg_classes(gsap_ctx *c, const gsap_options &o) {
  c->rs.seg_reg_max = o.seg_reg_max ? o.seg_reg_max : 64;
  c->rs.seg_count_min = o.seg_count_min ? o.seg_count_min : 512;
}
int32_t resolve_tile(const gsap_options &o) { return o.tile_max_B > 0 ? o.tile_max_B : 2048; }

RngKey rng_key(const gsap_options &o) {
  return RngKey{(uint32_t)o.seed, (uint32_t)(o.seed >> 32), o.chain};
}

void rebuild(gsap_ctx *c, int32_t B) {
  int p = c->prof.begin(1, c->st);
  rebuild_blockmodel(c->g, c->lab.as<int32_t>(), B, c->bm, c->rs, c->dense_max_B, c->tile_max_B,
                     c->st);
  c->prof.end(p, c->st);
  HIPX(hipGetLastError());
}

// Whether the rebuild after batch b of a sweep defers the CSR/CSC (a sweep's batches before
// its last, at a B whose dense matrix the incremental update keeps, unless
// options.csr_every_batch).
bool defers_csr(const gsap_ctx *c, const gsap_options &o, int32_t B, int32_t b) {
  return b + 1 < o.num_batches && o.csr_every_batch == 0 && dense_update_ok(c->bm, c->rs, B);
}

// The rebuild after nodal batch b: the moved vertices' edges moved between cells of the
// dense matrix when it holds the matrix of the labels before the batch, else a full rebuild.
// Without the dense matrix and with options.sparse_update = 1, the sparse update merges the
// batch's delta into the current CSR/CSC (a captured sweep of an odd number of batches would
// end on the other buffer set, so it rebuilds instead).
//
// defer_csr (from defers_csr: the batch's apply moved the block degrees): only the dense
// matrix follows the batch; the next batch samples from the dense matrix and the sweep's
// last batch rebuilds the CSR/CSC that the description length and the merges read.
void rebuild_after_batch(gsap_ctx *c, int32_t B, const gsap_options &o, int32_t b,
                         bool defer_csr = false) {
  int p = c->prof.begin(1, c->st);
  const int64_t BS = c->N / o.num_batches, lo = (int64_t)b * BS;
  const int32_t *lab = c->lab.as<int32_t>(), *olab = c->olab.as<int32_t>();
  c->rs.sparse_on = o.sparse_update == 1 && !(c->capturing && (o.num_batches & 1));
  const int64_t bound = b < (int32_t)c->delta_bound.size() ? c->delta_bound[b] : INT64_MAX;
  if (defer_csr && !dense_update_ok(c->bm, c->rs, B))
    throw Err(GSAP_ERR_INTERNAL, "deferred block-matrix update without the dense matrix");
  if (!rebuild_after_moves(c->g, lab, olab, lo, BS, B, c->bm, c->rs, c->st, defer_csr) &&
      !rebuild_after_moves_sparse(c->g, lab, olab, lo, BS, B, bound, c->bm, c->rs, c->st))
    rebuild_blockmodel(c->g, lab, B, c->bm, c->rs, c->dense_max_B, c->tile_max_B, c->st);
  c->prof.end(p, c->st);
  HIPX(hipGetLastError());
}

// Waits for all work queued on the stream and reads res[0..n) (device memory) and the block
// matrix's entry count: a one-thread kernel writes them into coherent host memory followed
// by a sequence number, which the host polls (no interrupt-driven wake-up between a sweep
// and the next).  The stream is queried now and then, so a failed launch ends the wait with
// its error.
// Readers of the CSR/CSC (the entry count, the description length, the merges) need the
// current labels' matrix: no deferred update may be pending.
void require_csr(const gsap_ctx *c) {
  if (c->bm.csr_stale)
    throw Err(GSAP_ERR_INTERNAL, "block-matrix CSR read while a deferred update is pending");
}

uint32_t publish(gsap_ctx *c, const double *res, int32_t n) {
  require_csr(c);
  const uint32_t seq = ++c->pub_seq;
  launch_publish(res, n, c->bm.rowptr + c->bm.B, c->pub_dev->v, &c->pub_dev->nnz,
                 &c->pub_dev->seq, seq, c->st);
  HIPX(hipGetLastError());
  return seq;
}

// Waits for publication seq; work queued after it may still run.
const HostPub &wait_published(gsap_ctx *c, uint32_t seq) {
  const uint32_t *f = &c->pub->seq;
  double next_query = now() + 1e-3;
  for (uint32_t it = 1; __atomic_load_n(f, __ATOMIC_ACQUIRE) != seq; ++it) {
    // the runtime is asked about the stream at most once per millisecond of waiting, so
    // that it does not slow the work queued behind the publication
    if ((it & 4095u) != 0 || now() < next_query) continue;
    next_query = now() + 1e-3;
    const hipError_t e = hipStreamQuery(c->st);
    if (e == hipErrorNotReady) continue;
    HIPX(e);
    // the stream has drained, so the stores are complete; allow for their visibility
    for (int k = 0; k < (1 << 20) && __atomic_load_n(f, __ATOMIC_ACQUIRE) != seq; ++k) {}
    if (__atomic_load_n(f, __ATOMIC_ACQUIRE) != seq)
      throw Err(GSAP_ERR_DEVICE, "published results did not reach host memory");
  }
  // per-kernel-class event timing reads its events after a stream synchronisation
  if (c->prof.on) HIPX(hipStreamSynchronize(c->st));
  return *c->pub;
}

const HostPub &publish_wait(gsap_ctx *c, const double *res, int32_t n) {
  return wait_published(c, publish(c, res, n));
}

int32_t read_nnz(gsap_ctx *c) { return publish_wait(c, nullptr, 0).nnz; }

int32_t avg_row(int64_t nnz, int32_t B) { return (int32_t)(B > 0 ? (nnz + B - 1) / B : 0); }

// total description length from the canonical data sum (partition.hpp:665-704, FP64)
double model_S(double data_S, int64_t N, int64_t E, int32_t B, int32_t model_kind) {
  const double LN2 = 0.6931471805599453094;
  double b = (double)B, dE = (double)E;
  double x = (b * b) / dE;
  double model;
  if (model_kind == GSAP_MODEL_CORRECTED)
    model = dE * ((1.0 + x) * (clog2_host(1.0 + x) * LN2) - x * (clog2_host(x) * LN2)) +
            (double)N * (clog2_host(b) * LN2);
  else
    model = dE * (1.0 + x) * (clog2_host(1.0 + x) * LN2) - x * (clog2_host(x) * LN2) +
            (double)N * (clog2_host(b) * LN2);
  return model - data_S;
}

// tile-sum region of nodal batch b (after the MDL's)
int64_t batch_tiles_off(const gsap_ctx *c) { return std::max<int64_t>(c->E, 1) / 1024 + 2; }
int64_t batch_tiles_stride(const gsap_ctx *c, int32_t nb) { return c->N / nb / 1024 + 1; }

void launch_mdl(gsap_ctx *c, double *out, uint32_t *ctr = nullptr, uint32_t add = 0) {
  require_csr(c);
  int p = c->prof.begin(4, c->st);
  BMView m = c->bm.view();
  launch_mdl_sum(m, c->bm.erow, c->E, c->tiles.as<double>(), out, ctr, add, c->st);
  c->prof.end(p, c->st);
}

// the sweep's closing sums: the description length's data term into res[0], each batch's
// deferred delta-S total into res[1 + b]; the batch counter advanced for the next sweep
void sweep_sums(gsap_ctx *c, const gsap_options &o, double *res, uint32_t *ctr) {
  require_csr(c);
  int p = c->prof.begin(4, c->st);
  BMView m = c->bm.view();
  launch_sweep_sums(m, c->bm.erow, c->E, c->tiles.as<double>(), batch_tiles_off(c),
                    batch_tiles_stride(c, o.num_batches), o.num_batches, c->N / o.num_batches,
                    res, ctr, (uint32_t)o.num_batches, c->st);
  c->prof.end(p, c->st);
}

double mdl_now(gsap_ctx *c, int32_t model_kind) {
  double *res = c->res.as<double>();
  launch_mdl(c, res);
  double h = 0;
  HIPX(hipMemcpyAsync(&h, res, 8, hipMemcpyDeviceToHost, c->st));
  HIPX(hipStreamSynchronize(c->st));
  return model_S(h, c->N, c->E, c->bm.B, model_kind);
}

int32_t find_root(int32_t *par, int32_t x) {
  int32_t r = x;
  while (par[r] != r) r = par[r];
  while (par[x] != r) { int32_t n = par[x]; par[x] = r; x = n; }
  return r;
}

// One merge phase: proposals + delta-S + argmin on the device, greedy merge order on the
// host (carry_out_best_merges, partition.hpp:26-78, as union-find), relabel on the device.
int32_t merge_phase(gsap_ctx *c, const gsap_options &o, int32_t B_to_merge, uint32_t phase,
                    std::vector<int32_t> *prop_out, std::vector<double> *dS_out,
                    std::vector<int32_t> *best_out, std::vector<double> *dSb_out) {
  const int32_t B = c->bm.B, np = o.num_proposals;
  hipStream_t st = c->st;
  BMView m = c->bm.view();
  const int64_t nnz = read_nnz(c);
  int32_t ar = avg_row(nnz, B);
  int p = c->prof.begin(2, st);
  launch_block_terms(m, c->OUT.as<double>(), nullptr, nullptr, rng_key(o), PhaseRef{nullptr, phase},
                     false, ar, st);
  c->prof.end(p, st);
  p = c->prof.begin(3, st);
  launch_merge(m, c->OUT.as<double>(), rng_key(o), phase, np, c->prop.as<int32_t>(),
               c->dS.as<double>(), nnz, o.delta_mode == GSAP_DELTA_EXACT, c->sof.as<int32_t>(),
               c->ctr.as<uint32_t>() + 1, st);
  c->prof.end(p, st);
  launch_argmin(c->dS.as<double>(), c->prop.as<int32_t>(), B, np, c->best.as<int32_t>(),
                c->dSbest.as<double>(), st);
  HIPX(hipGetLastError());
  // stable ascending order of the best delta-S (thrust::sort_by_key, partition.hpp:442-445):
  // a stable LSD radix sort of the order-preserving 64-bit keys on the device for large B,
  // std::stable_sort of the same keys on the host otherwise (identical order)
  const bool dev_sort = B >= 4096;
  // pinned staging sized for the graph (B <= N): delta-S, best partner, order, merge forest
  c->stage.ensure((size_t)c->N * 20);
  double *dsb = c->stage.as<double>();
  int32_t *best = (int32_t *)(dsb + B), *order = best + B, *map = order + B;
  HIPX(hipMemcpyAsync(best, c->best.p, 4 * (size_t)B, hipMemcpyDeviceToHost, st));
  if (dev_sort) {
    uint64_t *k0 = c->mk0.as<uint64_t>(), *k1 = c->mk1.as<uint64_t>();
    uint32_t *v0 = c->mv0.as<uint32_t>(), *v1 = c->mv1.as<uint32_t>();
    launch_order_keys(c->dSbest.as<double>(), B, k0, v0, st);
    radix_sort_pairs<uint64_t>(k0, v0, k1, v1, B, 64, c->tmp.as<uint32_t>(), st);
    HIPX(hipGetLastError());
    HIPX(hipMemcpyAsync(order, v0, 4 * (size_t)B, hipMemcpyDeviceToHost, st));
  }
  if (!dev_sort || dSb_out)
    HIPX(hipMemcpyAsync(dsb, c->dSbest.p, 8 * (size_t)B, hipMemcpyDeviceToHost, st));
  if (prop_out) {
    prop_out->resize((size_t)B * np);
    dS_out->resize((size_t)B * np);
    HIPX(hipMemcpyAsync(prop_out->data(), c->prop.p, 4 * (size_t)B * np, hipMemcpyDeviceToHost, st));
    HIPX(hipMemcpyAsync(dS_out->data(), c->dS.p, 8 * (size_t)B * np, hipMemcpyDeviceToHost, st));
  }
  publish_wait(c, nullptr, 0);
  c->prof.flush();
  if (best_out) best_out->assign(best, best + B);
  if (dSb_out) dSb_out->assign(dsb, dsb + B);
  if (!dev_sort) {
    std::vector<uint64_t> key(B);
    for (int32_t r = 0; r < B; ++r) key[r] = order_key(dsb[r]);
    std::iota(order, order + B, 0);
    std::stable_sort(order, order + B, [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  }
  // The merge loop links a block that is still a root (each block comes once in order) to
  // the root of another tree, so par stays a forest.
  int32_t *par = map;
  std::iota(par, par + B, 0);
  int32_t merged = 0, cidx = 0;
  while (merged < B_to_merge && cidx < B) {
    int32_t from = order[cidx++];
    int32_t to = find_root(par, best[from]);
    if (to != from) { par[from] = to; ++merged; }
  }
  // the device resolves every block's root, numbers the surviving roots in increasing id
  // order and relabels; the staging stays untouched until the next wait on this stream
  HIPX(hipMemcpyAsync(c->map.p, par, 4 * (size_t)B, hipMemcpyHostToDevice, st));
  launch_merge_relabel(c->map.as<int32_t>(), B, c->mv0.as<uint32_t>(), c->tmp.as<uint32_t>(),
                       c->lab.as<int32_t>(), c->N, st);
  HIPX(hipGetLastError());
  return B - merged;
}

// The shared neighbour samples of a nodal batch (its block terms); they only read the block
// matrix and the batch counter, and clear the proposal pass's class counters.
void batch_terms(gsap_ctx *c, const gsap_options &o, PhaseRef phase, int32_t ar) {
  int p = c->prof.begin(2, c->st);
  const BMBuf &bm = c->bm;
  if (bm.csr_stale)  // the samples from the dense matrix (and the log2 degrees with them)
    launch_block_terms_dense(bm.dense, bm.dt_B == bm.B ? c->rs.partials : nullptr,
                             bm.cs_B == bm.B ? c->rs.colsum : nullptr, bm.B,
                             bm.dout, bm.din, bm.ldout, bm.ldin, c->sof.as<int32_t>(),
                             c->coin.as<uint8_t>(), rng_key(o), phase, c->st,
                             c->dcnt.as<int32_t>());
  else
    launch_block_terms(bm.view(), nullptr, c->sof.as<int32_t>(), c->coin.as<uint8_t>(),
                       rng_key(o), phase, true, ar, c->st, c->dcnt.as<int32_t>());
  c->prof.end(p, c->st);
}

// The proposal pass's lists and counters (the counters cleared by the batch's block terms)
NodalPre nodal_pre(gsap_ctx *c) {
  NodalPre p;
  p.dlist = c->dlist.as<int4>();
  p.dcnt = c->dcnt.as<int32_t>();
  p.bounds = c->lists_bounds;
  // counted in profiled runs only
  if (c->prof.on) p.evstat = c->evstat.as<unsigned long long>();
  return p;
}

// One nodal batch on the current block matrix; the accepted delta-S sum lands in *dE_dev
// (nullptr: only the batch's tile sums, totalled by the sweep's launch_sweep_sums).
// terms_queued: the batch's block terms are already on the stream.
// defer_csr: the rebuild after this batch defers the CSR (rebuild_after_batch with the
// same flag), so the apply moves the block degrees.
void nodal_batch(gsap_ctx *c, const gsap_options &o, int32_t b, PhaseRef phase, int32_t ar,
                 double *dE_dev, bool terms_queued = false, bool defer_csr = false) {
  const int64_t BS = c->N / o.num_batches, lo = (int64_t)b * BS;
  hipStream_t st = c->st;
  BMView m = c->bm.view();
  if (!terms_queued) batch_terms(c, o, phase, ar);
  const int32_t *lists[NCLS];
  int32_t counts[NCLS];
  for (int k = 0; k < NCLS; ++k) {
    const int64_t off = c->list_off[(size_t)b * NCLS + k];
    const int64_t end = c->list_off[(size_t)b * NCLS + k + 1];
    lists[k] = c->vlist.as<int32_t>() + off;
    counts[k] = (int32_t)(end - off);
  }
  const bool mixed = o.precision == GSAP_PRECISION_MIXED;
  float *err = mixed ? c->errv.as<float>() : nullptr;
  uint8_t *evcls = mixed ? c->evcls.as<uint8_t>() : nullptr;
  // the labels as the last rebuild saw them (it made the 16-bit copy when B <= 65535)
  const LabView lv{c->lab.as<int32_t>(), c->bm.B <= 65535 ? c->rs.lab16 : nullptr};
  const NodalPre pre = nodal_pre(c);
  int p = -1;
  for (int stage : {NODAL_PROPOSE, NODAL_MOVES}) {
    if (stage == NODAL_PROPOSE && c->g.wmax != 1) continue;  // unit-weight graphs only
    p = c->prof.begin(stage == NODAL_PROPOSE ? 6 : 0, st);
    launch_nodal(c->g, lv, m, c->sof.as<int32_t>(), c->coin.as<uint8_t>(),
                 rng_key(o), phase, lo, BS, lists, counts, c->big_lds[b], c->prop.as<int32_t>(),
                 c->dS.as<double>(), err, evcls, o.delta_mode == GSAP_DELTA_EXACT,
                 o.accept == GSAP_ACCEPT_MH, o.beta, pre, stage, st);
    if (stage == NODAL_PROPOSE) c->prof.end(p, st);
  }
  if (mixed) {
    uint32_t *ms = c->mixstat.as<uint32_t>();
    launch_mix_stats(err, evcls, BS, ms, reinterpret_cast<unsigned long long *>(ms + 2), st);
  }
  c->prof.end(p, st);
  p = c->prof.begin(5, st);
  // the 16-bit label copy follows the batch's changes (the rebuild after the batch then
  // finds it current)
  DegMoves dm;
  if (defer_csr) {
    const bool unit = c->g.wmax == 1;
    dm.optr = c->g.optr; dm.iptr = c->g.iptr;
    dm.kout = unit ? nullptr : c->g.kout; dm.kin = unit ? nullptr : c->g.kin;
    dm.B = c->bm.B; dm.dout = c->bm.dout; dm.din = c->bm.din;
  }
  launch_apply_sum(c->lab.as<int32_t>(), c->prop.as<int32_t>(), c->dS.as<double>(), lo, BS,
                   o.accept == GSAP_ACCEPT_MH, c->olab.as<int32_t>(),
                   c->bm.B <= 65535 ? c->rs.lab16 : nullptr,
                   c->tiles.as<double>() + batch_tiles_off(c) +
                       (int64_t)b * batch_tiles_stride(c, o.num_batches),
                   dE_dev, dm, st);
  c->prof.end(p, st);
  HIPX(hipGetLastError());
}

void check_options(const gsap_options &o) {
  if (o.num_proposals < 1 || o.num_proposals > 32)
    throw Err(GSAP_ERR_INVALID, "num_proposals must be in [1, 32]");
  if (o.num_batches < 1 || o.num_batches > 63)
    throw Err(GSAP_ERR_INVALID, "num_batches must be in [1, 63]");
  if (o.max_nodal_itr < 1) throw Err(GSAP_ERR_INVALID, "max_nodal_itr must be >= 1");
  if (o.moving_avg_window < 2) throw Err(GSAP_ERR_INVALID, "moving_avg_window must be >= 2");
  if (o.precision != GSAP_PRECISION_FP64 && o.precision != GSAP_PRECISION_MIXED)
    throw Err(GSAP_ERR_INVALID, "unknown precision");
  if (o.model != GSAP_MODEL_REFERENCE && o.model != GSAP_MODEL_CORRECTED)
    throw Err(GSAP_ERR_INVALID, "unknown model");
  if (o.dense_max_B < 0 || o.tile_max_B < 0)
    throw Err(GSAP_ERR_INVALID, "dense_max_B and tile_max_B must be >= 0");
  {
    const NodalBounds b = resolve_nodal(o);
    if (b.max16 < -1 || b.max16 > 16 || b.max32 < -1 || b.max32 > 32 || b.maxwave < -1 ||
        b.maxwave > 128)
      throw Err(GSAP_ERR_INVALID, "nodal_max16, nodal_max32, nodal_maxwave must be -1, 0 or at "
                                  "most 16, 32 and 128");
  }
  if (o.seg_reg_max != 0 && o.seg_reg_max != 16 && o.seg_reg_max != 32 && o.seg_reg_max != 64)
    throw Err(GSAP_ERR_INVALID, "seg_reg_max must be 0, 16, 32 or 64");
  if (o.seg_count_min != 0 && (o.seg_count_min < 512 || o.seg_count_min > 4096))
    throw Err(GSAP_ERR_INVALID, "seg_count_min must be 0 or in 512..4096");
  if (o.sparse_update != 0 && o.sparse_update != 1)
    throw Err(GSAP_ERR_INVALID, "sparse_update must be 0 or 1");
  if (o.csr_every_batch != 0 && o.csr_every_batch != 1)
    throw Err(GSAP_ERR_INVALID, "csr_every_batch must be 0 or 1");

  if (o.delta_mode != GSAP_DELTA_REFERENCE && o.delta_mode != GSAP_DELTA_EXACT)
    throw Err(GSAP_ERR_INVALID, "unknown delta_mode");
  if (o.accept != GSAP_ACCEPT_GREEDY && o.accept != GSAP_ACCEPT_MH)
    throw Err(GSAP_ERR_INVALID, "unknown accept");
  if (o.accept == GSAP_ACCEPT_MH && !(o.beta >= 0.0 && o.beta <= 1e300))
    throw Err(GSAP_ERR_INVALID, "beta must be finite and >= 0");
  // the mixed form's FP32 pass proves rejections of dS >= 0, which MH may accept
  if (o.accept == GSAP_ACCEPT_MH && o.precision == GSAP_PRECISION_MIXED)
    throw Err(GSAP_ERR_INVALID, "Metropolis-Hastings acceptance needs precision FP64");
}

// A level's captured sweep: the executable graph, destroyed when the level ends or an error
// unwinds.
struct GraphGuard {
  gsap_ctx *c;
  hipGraphExec_t exec = nullptr;
  explicit GraphGuard(gsap_ctx *ctx) : c(ctx) {}
  template <class F>
  void capture(F &&body) {
    HIPX(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    c->capturing = true;
    try {
      body();
      c->capturing = false;
    } catch (...) {
      c->capturing = false;
      (void)hipStreamEndCapture(c->st, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPX(hipStreamEndCapture(c->st, &g));
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      exec = nullptr;
      throw Err(GSAP_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
  }
  ~GraphGuard() {
    if (exec) {
      (void)hipStreamSynchronize(c->st);
      (void)hipGraphExecDestroy(exec);
    }
  }
};


void partition_impl(gsap_ctx *c, const gsap_graph_i32 *gin, const gsap_options &o, int32_t *out,
                    gsap_report *rep) {
  check_options(o);
  const double t0 = now();
  set_device(c);
  // gin == nullptr: the graph a previous upload left on the device
  if (gin) upload_graph(c, gin);
  else if (!c->resident) throw Err(GSAP_ERR_INVALID, "no graph uploaded to this context");
  const gsap_graph_i32 hp = resident_ptrs(c);
  build_batch_lists(c, &hp, o.num_batches, resolve_nodal(o));
  alloc_work(c, o.num_proposals, resolve_dense(o));
  apply_seg_classes(c, o);
  c->tile_max_B = resolve_tile(o);
  const double t_upload = now() - t0;
  c->trace.clear();
  c->prof.reset(o.profile != 0);
  c->prof.mask = o.profile_classes ? (uint32_t)o.profile_classes : 0xFFu;
  const bool tr = o.trace != 0;
  const int64_t N = c->N;
  hipStream_t st = c->st;
  if (c->E == 0) {
    // no edges: the model term divides by E (partition.hpp:696-699); one block, S = 0
    std::fill(out, out + N, 0);
    if (rep) {
      std::memset(rep, 0, sizeof(*rep));
      rep->B = 1;
      rep->t_total = now() - t0;
      rep->t_upload = t_upload;
    }
    return;
  }
  int32_t *lab = c->lab.as<int32_t>();
  double *res = c->res.as<double>();
  launch_iota(lab, N, st);
  HIPX(hipMemsetAsync(c->mixstat.p, 0, 32, st));
  HIPX(hipMemsetAsync(c->evstat.p, 0, 16 * EV_SLOTS, st));
  uint32_t *ctr = c->ctr.as<uint32_t>();
  HIPX(hipMemsetAsync(ctr, 0, 4, st));
  int32_t B = (int32_t)N;
  int32_t B_to_merge = (int32_t)((float)B * o.reduction_rate);
  rebuild(c, B);
  gsap_golden_slot slots[3];  // large, med, small; slot k owns label snapshot k
  gsap_golden_init(slots);
  uint32_t merge_ctr = 0, batch_ctr = 0;
  int32_t levels = 0, sweeps = 0;
  int64_t batches = 0;
  double S = 0.0, t_merge = 0.0, t_sweep = 0.0;
  bool optimal = false;
  const int64_t BS = N / o.num_batches;
  std::vector<int32_t> hprop, hlab, hbefore, hafter;
  std::vector<double> hdS;
  auto rec = [&](int32_t kind, int32_t Bv, int32_t a, int32_t b, uint64_t h1, uint64_t h2,
                 double x, double y) {
    c->trace.push_back(gsap_trace_rec{kind, Bv, a, b, h1, h2, x, y});
  };
  auto labels_hash = [&]() {
    hlab.resize(N);
    HIPX(hipMemcpyAsync(hlab.data(), lab, 4 * (size_t)N, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    return fnv(hlab.data(), 4 * N);
  };
  while (!optimal) {
    const double tm = now();
    int32_t nB = merge_phase(c, o, B_to_merge, merge_ctr, tr ? &hprop : nullptr,
                             tr ? &hdS : nullptr, nullptr, nullptr);
    if (tr) rec(1, B, nB, (int32_t)merge_ctr, fnv(hprop.data(), 4 * (int64_t)hprop.size()),
                labels_hash(), 0.0, 0.0);
    ++merge_ctr;
    B = nB;
    rebuild(c, B);
    int32_t ar = avg_row(read_nnz(c), B);
    t_merge += now() - tm;
    std::vector<double> dE(o.max_nodal_itr, 0.0);
    const int64_t batches_before = batches;
    // one sweep: num_batches x (nodal batch + rebuild), the description length, and the
    // device batch counter advanced; with options.graph (and no tracing or event timing:
    // HIP cannot time events recorded inside a graph) it is captured once per level (B and
    // every launch shape are fixed within a level) and replayed as a HIP graph
    const bool use_graph = o.graph && !tr && !c->prof.on;
    GraphGuard gg(c);
    const int32_t ar_level = ar;
    bool terms_queued = false;  // the sweep's first block terms went ahead of the last wait
    for (int itr = 0; itr < o.max_nodal_itr; ++itr) {
      const double ts = now();
      if (use_graph) {
        if (!gg.exec) gg.capture([&] {
          for (int b = 0; b < o.num_batches; ++b) {
            const bool defer = defers_csr(c, o, B, b);
            nodal_batch(c, o, b, PhaseRef{ctr, (uint32_t)b}, ar_level, nullptr, false, defer);
            rebuild_after_batch(c, B, o, b, defer);
          }
          sweep_sums(c, o, res, ctr);
        });
        HIPX(hipGraphLaunch(gg.exec, st));
        batch_ctr += (uint32_t)o.num_batches;
        batches += o.num_batches;
      }
      for (int b = 0; !use_graph && b < o.num_batches; ++b) {
        const int32_t *lab_b = c->lab.as<int32_t>() + (int64_t)b * BS;
        if (tr) {
          hbefore.resize(BS);
          HIPX(hipMemcpyAsync(hbefore.data(), lab_b, 4 * (size_t)BS, hipMemcpyDeviceToHost, st));
        }
        // traced: every batch's sum at once (read below); else totalled at the sweep's end
        const bool defer = defers_csr(c, o, B, b);
        nodal_batch(c, o, b, PhaseRef{ctr, (uint32_t)b}, ar, tr ? res + 1 + b : nullptr,
                    b == 0 && terms_queued, defer);
        rebuild_after_batch(c, B, o, b, defer);
        if (tr) {
          double e = 0;
          hprop.resize(BS);
          hafter.resize(BS);
          HIPX(hipMemcpyAsync(hprop.data(), c->prop.p, 4 * (size_t)BS, hipMemcpyDeviceToHost, st));
          HIPX(hipMemcpyAsync(hafter.data(), lab_b, 4 * (size_t)BS, hipMemcpyDeviceToHost, st));
          HIPX(hipMemcpyAsync(&e, res + 1 + b, 8, hipMemcpyDeviceToHost, st));
          HIPX(hipStreamSynchronize(st));
          int64_t moved = 0;  // vertices of the batch whose label changed
          for (int64_t i = 0; i < BS; ++i) moved += hafter[i] != hbefore[i];
          rec(2, B, itr, b, fnv(hprop.data(), 4 * BS), labels_hash(), e, (double)moved);
        }
        ++batch_ctr;
        ++batches;
      }
      if (!use_graph) {
        if (tr) launch_mdl(c, res, ctr, (uint32_t)o.num_batches);
        else sweep_sums(c, o, res, ctr);
      }
      double h[64];
      const uint32_t seq = publish(c, res, 1 + o.num_batches);
      // The next sweep's first block terms are queued behind the publication, so the device
      // works while the host takes the convergence decision; they read only the block matrix
      // and the advanced counter (no state changes; unused when the level ends here).  Their
      // lane-group width follows the previous sweep's average row, which does not change
      // what they compute.
      terms_queued = !use_graph && !tr && itr + 1 < o.max_nodal_itr;
      if (terms_queued) batch_terms(c, o, PhaseRef{ctr, 0}, ar);
      const HostPub &hp = wait_published(c, seq);
      std::memcpy(h, hp.v, 8 * (size_t)(1 + o.num_batches));
      const int32_t nnz = hp.nnz;
      c->prof.flush();
      ar = avg_row(nnz, B);
      for (int b = 0; b < o.num_batches; ++b) dE[itr] = dE[itr] + h[1 + b];
      S = model_S(h[0], N, c->E, B, o.model);
      ++sweeps;
      t_sweep += now() - ts;
      if (tr) {
        std::vector<int32_t> col(nnz), val(nnz);
        HIPX(hipMemcpyAsync(col.data(), c->bm.col, 4 * (size_t)nnz, hipMemcpyDeviceToHost, st));
        HIPX(hipMemcpyAsync(val.data(), c->bm.val, 4 * (size_t)nnz, hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
        rec(3, B, itr, 0, fnv(col.data(), 4 * (int64_t)nnz), fnv(val.data(), 4 * (int64_t)nnz), S,
            dE[itr]);
      }
      const int32_t bracketed = std::isfinite(slots[0].S) && std::isfinite(slots[1].S) &&
                                std::isfinite(slots[2].S);
      if (gsap_converged(dE.data(), itr, o.moving_avg_window, S, bracketed, o.threshold1,
                         o.threshold2))
        break;
    }
    if (c->level_rec) {  // the autotuner's snapshot of the level's labelling
      LevelSnap ls;
      ls.B = B;
      ls.batches = batches - batches_before;
      ls.lab.resize((size_t)N);
      HIPX(hipMemcpyAsync(ls.lab.data(), lab, 4 * (size_t)N, hipMemcpyDeviceToHost, st));
      HIPX(hipStreamSynchronize(st));
      c->level_rec->push_back(std::move(ls));
    }
    // golden section (partition.hpp:80-322); snapshots hold labels only, the block
    // matrix of a restored snapshot is rebuilt from them
    const gsap_golden_slot P{B, B_to_merge, S, -1, 0};
    gsap_golden_slot next;
    int32_t store = 0, restore = 0, opt = 0;
    gsap_golden_step(slots, &P, o.reduction_rate, &store, &restore, &next, &opt);
    HIPX(hipMemcpyAsync(c->snap[store].p, lab, 4 * (size_t)N, hipMemcpyDeviceToDevice, st));
    optimal = opt != 0;
    B = next.B;
    B_to_merge = next.B_to_merge;
    S = next.S;
    HIPX(hipMemcpyAsync(lab, c->snap[restore].p, 4 * (size_t)N, hipMemcpyDeviceToDevice, st));
    rebuild(c, B);
    ++levels;
    if (tr) rec(4, B, B_to_merge, optimal ? 1 : 0, labels_hash(), 0, S, 0.0);
    if (o.max_levels > 0 && levels >= o.max_levels) break;
  }
  HIPX(hipMemcpyAsync(out, lab, 4 * (size_t)N, hipMemcpyDeviceToHost, st));
  uint32_t mix[8] = {0};
  HIPX(hipMemcpyAsync(mix, c->mixstat.p, 32, hipMemcpyDeviceToHost, st));
  uint64_t ev[2 * EV_SLOTS] = {0};
  HIPX(hipMemcpyAsync(ev, c->evstat.p, sizeof(ev), hipMemcpyDeviceToHost, st));
  HIPX(hipStreamSynchronize(st));
  c->prof.flush();
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->B = B; rep->levels = levels; rep->merge_phases = (int32_t)merge_ctr;
    rep->sweeps = sweeps; rep->batches = batches; rep->S = S;
    rep->t_total = now() - t0; rep->t_merge = t_merge; rep->t_sweep = t_sweep;
    rep->t_upload = t_upload;
    rep->sweep_edges = c->E * (int64_t)sweeps;
    rep->n_trace = (int64_t)c->trace.size();
    for (int k = 0; k < 8; ++k) { rep->kernel_ms[k] = c->prof.ms[k]; rep->kernel_calls[k] = c->prof.calls[k]; }
    if (c->prof.on) rep->t_rebuild = c->prof.ms[1] / 1e3;  // class 1: block-matrix rebuilds
    float mx;
    std::memcpy(&mx, &mix[0], 4);
    uint64_t cnt[2];
    std::memcpy(cnt, &mix[2], 16);
    rep->max_abs_err_fp32 = (double)mx;
    rep->rechecks = (int64_t)cnt[0];
    rep->fp64_evals = (int64_t)cnt[1];
    int64_t evals = 0, adj = 0;
    for (int k = 0; k < EV_SLOTS; ++k) {
      evals += (int64_t)ev[2 * k];
      adj += (int64_t)ev[2 * k + 1];
    }
    const bool counted = c->prof.on && c->g.wmax == 1;
    rep->nodal_evals = counted ? evals : -1;
    rep->nodal_eval_adj = counted ? adj : -1;
  }
}

gsap_options defaults() {
  gsap_options o;
  std::memset(&o, 0, sizeof(o));
  o.seed = 1;
  o.num_proposals = 10;
  o.reduction_rate = 0.35f;
  o.max_nodal_itr = 100;
  o.num_batches = 4;
  o.threshold1 = 5e-4;
  o.threshold2 = 1e-4;
  o.moving_avg_window = 3;
  o.max_levels = 0;
  o.precision = GSAP_PRECISION_FP64;
  o.beta = 3.0;
  o.chain = 0;
  o.trace = 0;
  o.dense_max_B = 0;
  o.profile = 0;
  o.model = GSAP_MODEL_REFERENCE;
  o.workers = 0;
  return o;
}

template <class F>
int guarded(gsap_ctx *c, F &&f) {
  if (!c) return GSAP_ERR_INVALID;
  try {
    c->err.clear();
    f();
    return GSAP_OK;
  } catch (const Err &e) {
    c->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    c->err = e.what();
    return GSAP_ERR_INTERNAL;
  }
}

void setup_unit(gsap_ctx *c, const gsap_graph_i32 *g, const gsap_options *o, const int32_t *labels,
                int32_t B) {
  gsap_options oo = o ? *o : defaults();
  check_options(oo);
  set_device(c);
  upload_graph(c, g);
  build_batch_lists(c, g, oo.num_batches, resolve_nodal(oo));
  alloc_work(c, oo.num_proposals, resolve_dense(oo));
  apply_seg_classes(c, oo);
  c->tile_max_B = resolve_tile(oo);
  if (B <= 0 || B > c->N) throw Err(GSAP_ERR_INVALID, "B must be in [1, N]");
  for (int64_t v = 0; v < c->N; ++v)
    if (labels[v] < 0 || labels[v] >= B) throw Err(GSAP_ERR_INVALID, "label out of range");
  HIPX(hipMemcpyAsync(c->lab.p, labels, 4 * (size_t)c->N, hipMemcpyHostToDevice, c->st));
  rebuild(c, B);
}

// Runs jobs 0..n-1 on up to `want` worker sub-contexts, one host thread each.
template <class F>
void run_pool(gsap_ctx *c, int32_t n, int32_t want, F &&job) {
  const int32_t W = std::max(1, std::min(n, want > 0 ? want : 8));
  while ((int32_t)c->workers.size() < W) {
    gsap_ctx *w = nullptr;
    if (gsap_create(&w, c->device) != GSAP_OK)
      throw Err(GSAP_ERR_DEVICE, "could not create a worker context");
    c->workers.push_back(w);
  }
  std::atomic<int32_t> next(0);
  std::vector<std::string> errs(n);
  std::vector<int> codes(n, GSAP_OK);
  std::vector<std::thread> th;
  for (int32_t w = 0; w < W; ++w)
    th.emplace_back([&, w] {
      for (int32_t k; (k = next++) < n;) {
        codes[k] = job(c->workers[w], k);
        if (codes[k] != GSAP_OK) errs[k] = gsap_last_error(c->workers[w]);
      }
    });
  for (auto &t : th) t.join();
  for (int32_t k = 0; k < n; ++k)
    if (codes[k] != GSAP_OK)
      throw Err(codes[k], "job " + std::to_string(k) + ": " + errs[k]);
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

void gsap_options_default(gsap_options *o) {
  if (o) *o = defaults();
}

int gsap_create(gsap_ctx **ctx, int device) {
  if (!ctx) return GSAP_ERR_INVALID;
  *ctx = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GSAP_ERR_DEVICE;
  if (device < 0 || device >= n) return GSAP_ERR_INVALID;
  gsap_ctx *c = new gsap_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess ||
      hipHostMalloc((void **)&c->pub, sizeof(HostPub), hipHostMallocMapped | hipHostMallocCoherent) !=
          hipSuccess ||
      hipHostGetDevicePointer((void **)&c->pub_dev, c->pub, 0) != hipSuccess) {
    if (c->pub) (void)hipHostFree(c->pub);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return GSAP_ERR_DEVICE;
  }
  std::memset(c->pub, 0, sizeof(HostPub));
  // this device's address of the no-allocation entry point's scratch array
  if (!noalloc_init_device(device)) {
    (void)hipHostFree(c->pub);
    (void)hipStreamDestroy(c->st);
    delete c;
    return GSAP_ERR_DEVICE;
  }
  *ctx = c;
  return GSAP_OK;
}

int gsap_destroy(gsap_ctx *c) {
  if (!c) return GSAP_OK;
  for (gsap_ctx *w : c->workers) gsap_destroy(w);
  c->workers.clear();
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->st);
  DBuf *all[] = {&c->optr, &c->oadj, &c->ow, &c->iptr, &c->iadj, &c->iw, &c->kout, &c->kin,
                 &c->esrc, &c->idst, &c->sloop, &c->hoff, &c->hbo, &c->hbi, &c->hpool, &c->lab,
                 &c->dlist, &c->dcnt, &c->snap[0],
                 &c->snap[1], &c->snap[2], &c->rowptr, &c->colptr, &c->col, &c->val, &c->erow,
                 &c->row, &c->cval, &c->dout, &c->din, &c->ldout, &c->ldin, &c->dense,
                 &c->gcol, &c->grow, &c->rowptr2, &c->colptr2, &c->col2, &c->val2,
                 &c->erow2, &c->row2, &c->cval2, &c->gcol2, &c->grow2, &c->rk, &c->rw, &c->mwork,
                 &c->k0,
                 &c->k1, &c->tmp, &c->partials, &c->soff, &c->su, &c->suoff,
                 &c->lists, &c->ncl, &c->cntc, &c->vrel, &c->chunk_seg, &c->lab16, &c->colsum, &c->vlist,
                 &c->OUT, &c->sof, &c->coin, &c->prop, &c->dS, &c->best, &c->dSbest, &c->map,
                 &c->tiles, &c->res, &c->ctr, &c->mk0, &c->mk1, &c->mv0,
                 &c->mv1, &c->errv, &c->evcls, &c->mixstat, &c->evstat, &c->olab};
  for (DBuf *b : all) b->release();
  try { c->prof.destroy(); } catch (...) {}
  c->stage.release();
  (void)hipHostFree(c->pub);
  (void)hipStreamDestroy(c->st);
  delete c;
  return GSAP_OK;
}

const char *gsap_last_error(const gsap_ctx *c) { return c ? c->err.c_str() : "no context"; }

int gsap_partition_i32(gsap_ctx *c, const gsap_graph_i32 *g, const gsap_options *o,
                       int32_t *partitions_out, gsap_report *report) {
  return guarded(c, [&] {
    if (!partitions_out) throw Err(GSAP_ERR_INVALID, "partitions_out is NULL");
    gsap_options oo = o ? *o : defaults();
    partition_impl(c, g, oo, partitions_out, report);
  });
}

int gsap_upload_graph_i32(gsap_ctx *c, const gsap_graph_i32 *g) {
  return guarded(c, [&] {
    if (!g) throw Err(GSAP_ERR_INVALID, "graph is NULL");
    set_device(c);
    upload_graph(c, g);
  });
}

int gsap_partition_resident_i32(gsap_ctx *c, const gsap_options *o, int64_t N,
                                int32_t *partitions_out, gsap_report *report) {
  return guarded(c, [&] {
    if (!partitions_out) throw Err(GSAP_ERR_INVALID, "partitions_out is NULL");
    if (!c->resident) throw Err(GSAP_ERR_INVALID, "no graph uploaded to this context");
    if (N != c->N)
      throw Err(GSAP_ERR_INVALID, "N (" + std::to_string(N) + ") is not the resident graph's " +
                                      std::to_string(c->N) + " vertices");
    gsap_options oo = o ? *o : defaults();
    partition_impl(c, nullptr, oo, partitions_out, report);
  });
}

int gsap_partition_batch_i32(gsap_ctx *c, const gsap_graph_i32 *graphs, int32_t n_graphs,
                             const gsap_options *o, int32_t *const *outs, gsap_report *reports) {
  return guarded(c, [&] {
    if (!graphs || !outs || n_graphs < 0) throw Err(GSAP_ERR_INVALID, "NULL argument");
    const gsap_options oo = o ? *o : defaults();
    check_options(oo);
    run_pool(c, n_graphs, oo.workers, [&](gsap_ctx *w, int32_t k) {
      return gsap_partition_i32(w, &graphs[k], &oo, outs[k], reports ? &reports[k] : nullptr);
    });
  });
}

int gsap_partition_chains_i32(gsap_ctx *c, const gsap_graph_i32 *g, const gsap_options *o,
                              int32_t n_chains, int32_t *out, gsap_report *report,
                              double *S_chain, int32_t *best_chain) {
  return guarded(c, [&] {
    if (!g || !out || n_chains < 1) throw Err(GSAP_ERR_INVALID, "bad argument");
    const gsap_options oo = o ? *o : defaults();
    check_options(oo);
    validate(g);
    std::vector<std::vector<int32_t>> labs(n_chains, std::vector<int32_t>((size_t)g->N));
    std::vector<gsap_report> reps(n_chains);
    run_pool(c, n_chains, oo.workers, [&](gsap_ctx *w, int32_t k) {
      gsap_options ok = oo;
      ok.chain = oo.chain + (uint32_t)k;
      return gsap_partition_i32(w, g, &ok, labs[k].data(), &reps[k]);
    });
    int32_t best = 0;
    for (int32_t k = 1; k < n_chains; ++k)
      if (reps[k].S < reps[best].S) best = k;
    std::copy(labs[best].begin(), labs[best].end(), out);
    if (report) *report = reps[best];
    if (S_chain)
      for (int32_t k = 0; k < n_chains; ++k) S_chain[k] = reps[k].S;
    if (best_chain) *best_chain = best;
  });
}

int gsap_autotune_i32(gsap_ctx *c, const gsap_graph_i32 *g, gsap_options *o) {
  return guarded(c, [&] {
    if (!o) throw Err(GSAP_ERR_INVALID, "options is NULL");
    gsap_options oo = *o;
    check_options(oo);
    set_device(c);
    // 1. one partition with the caller's options, keeping every level's labelling
    std::vector<LevelSnap> levels;
    {
      gsap_options po = oo;
      po.trace = 0;
      po.profile = 0;
      po.graph = 0;
      std::vector<int32_t> out((size_t)std::max<int64_t>(g->N, 1));
      c->level_rec = &levels;
      struct Unset {
        gsap_ctx *c;
        ~Unset() { c->level_rec = nullptr; }
      } unset{c};
      partition_impl(c, g, po, out.data(), nullptr);
    }
    if (levels.empty() || c->E == 0) return;
    // 2. up to 8 snapshots spread over the distinct block counts, each weighted by the
    //    batches run at the levels nearest to it in log B
    std::vector<const LevelSnap *> distinct;
    {
      std::vector<const LevelSnap *> byB;
      for (const LevelSnap &l : levels) byB.push_back(&l);
      std::stable_sort(byB.begin(), byB.end(),
                       [](const LevelSnap *x, const LevelSnap *y) { return x->B > y->B; });
      for (const LevelSnap *l : byB)
        if (distinct.empty() || distinct.back()->B != l->B) distinct.push_back(l);
    }
    const size_t K = std::min<size_t>(8, distinct.size());
    std::vector<const LevelSnap *> snaps;
    for (size_t k = 0; k < K; ++k)
      snaps.push_back(distinct[K == 1 ? 0 : k * (distinct.size() - 1) / (K - 1)]);
    std::vector<double> wt(snaps.size(), 0.0);
    for (const LevelSnap &l : levels) {
      size_t best = 0;
      for (size_t k = 1; k < snaps.size(); ++k)
        if (std::fabs(std::log((double)snaps[k]->B / l.B)) <
            std::fabs(std::log((double)snaps[best]->B / l.B)))
          best = k;
      wt[best] += (double)std::max<int64_t>(l.batches, 1);
    }
    const int64_t N = c->N;
    const int32_t dmax = (int32_t)std::min<int64_t>(16384, N);
    alloc_work(c, oo.num_proposals, dmax);
    hipStream_t st = c->st;
    hipEvent_t e0, e1;
    HIPX(hipEventCreate(&e0));
    HIPX(hipEventCreate(&e1));
    struct Ev {
      hipEvent_t a, b;
      ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
    } ev{e0, e1};
    auto elapsed = [&]() {
      HIPX(hipEventSynchronize(e1));
      float ms = 0;
      HIPX(hipEventElapsedTime(&ms, e0, e1));
      return (double)ms;
    };
    auto load = [&](const LevelSnap *sn) {
      HIPX(hipMemcpyAsync(c->lab.p, sn->lab.data(), 4 * (size_t)N, hipMemcpyHostToDevice, st));
    };
    const uint32_t *ctr = c->ctr.as<uint32_t>();
    HIPX(hipMemsetAsync(c->ctr.p, 0, 8, st));
    // device time of a rebuild of the snapshot's labelling (best of 3)
    auto t_rebuild = [&](const LevelSnap *sn) {
      load(sn);
      rebuild(c, sn->B);
      double best = 1e30;
      for (int rep = 0; rep < 3; ++rep) {
        HIPX(hipEventRecord(e0, st));
        rebuild(c, sn->B);
        HIPX(hipEventRecord(e1, st));
        best = std::min(best, elapsed());
      }
      return best;
    };
    // device time of a nodal batch from the snapshot (nodal kernels and the labels' apply;
    // with rebuild = true also the rebuild that follows it), best of 3 after a warm-up
    auto t_batch = [&](const LevelSnap *sn, bool with_rebuild) {
      double best = 1e30;
      for (int rep = 0; rep < 4; ++rep) {
        load(sn);
        rebuild(c, sn->B);
        const int32_t ar = avg_row(read_nnz(c), sn->B);
        HIPX(hipEventRecord(e0, st));
        nodal_batch(c, oo, 0, PhaseRef{ctr, 0}, ar, c->res.as<double>() + 1);
        if (with_rebuild) rebuild_after_batch(c, sn->B, oo, 0);
        HIPX(hipEventRecord(e1, st));
        const double t = elapsed();
        if (rep > 0) best = std::min(best, t);
      }
      return best;
    };
    // a candidate replaces the default only when it saves more than 1 % of the weighted time
    auto pick = [](const std::vector<double> &cost, size_t def) {
      size_t b = def;
      for (size_t k = 0; k < cost.size(); ++k)
        if (cost[k] < cost[b]) b = k;
      return cost[b] < 0.99 * cost[def] ? b : def;
    };
    int32_t tile = resolve_tile(oo), dense = resolve_dense(oo);
    // 3. row tile against segmented rebuild at the snapshots in (128, 8192] (dense both ways)
    {
      std::vector<int32_t> cand{tile, SMALL_B_TUNE};
      std::vector<double> tt, ts, ws;
      std::vector<int32_t> Bs;
      for (size_t k = 0; k < snaps.size(); ++k) {
        const int32_t B = snaps[k]->B;
        if (B <= SMALL_B_TUNE || B > 8192 || B > dmax) continue;
        c->dense_max_B = dmax;
        c->tile_max_B = B;
        tt.push_back(t_rebuild(snaps[k]));
        c->tile_max_B = 0;
        ts.push_back(t_rebuild(snaps[k]));
        ws.push_back(wt[k]);
        Bs.push_back(B);
        cand.push_back(B);
      }
      if (!Bs.empty()) {
        std::vector<double> cost;
        for (int32_t T : cand) {
          double x = 0;
          for (size_t j = 0; j < Bs.size(); ++j) x += ws[j] * (Bs[j] <= T ? tt[j] : ts[j]);
          cost.push_back(x);
        }
        tile = std::max<int32_t>(SMALL_B_TUNE, cand[pick(cost, 0)]);
      }
    }
    c->tile_max_B = tile;
    // 4. dense against CSR lookups over a whole batch at the snapshots in [2048, 16384]
    {
      std::vector<int32_t> cand{dense};
      std::vector<double> td, tc, ws;
      std::vector<int32_t> Bs;
      for (size_t k = 0; k < snaps.size(); ++k) {
        const int32_t B = snaps[k]->B;
        if (B < 2048 || B > dmax) continue;
        c->dense_max_B = B;
        td.push_back(t_batch(snaps[k], true));
        c->dense_max_B = B - 1;
        tc.push_back(t_batch(snaps[k], true));
        ws.push_back(wt[k]);
        Bs.push_back(B);
        cand.push_back(B);
        cand.push_back(B - 1);
      }
      if (!Bs.empty()) {
        std::vector<double> cost;
        for (int32_t D : cand) {
          double x = 0;
          for (size_t j = 0; j < Bs.size(); ++j) x += ws[j] * (Bs[j] <= D ? td[j] : tc[j]);
          cost.push_back(x);
        }
        dense = std::max<int32_t>(2047, cand[pick(cost, 0)]);
      }
    }
    c->dense_max_B = dense;
    // 5. segment-sort classes at the snapshots the segmented rebuild serves
    const int32_t seg_cand[][2] = {{64, 512}, {32, 512}, {16, 512}, {64, 4096}, {32, 4096}};
    {
      std::vector<double> cost(5, 0.0);
      bool any = false;
      for (size_t k = 0; k < snaps.size(); ++k) {
        const int32_t B = snaps[k]->B;
        const bool dn = B <= dense;
        if (B <= SMALL_B_TUNE || (dn && B <= tile)) continue;
        any = true;
        for (int j = 0; j < 5; ++j) {
          c->rs.seg_reg_max = seg_cand[j][0];
          c->rs.seg_count_min = seg_cand[j][1];
          cost[j] += wt[k] * t_rebuild(snaps[k]);
        }
      }
      const size_t j = any ? pick(cost, 0) : 0;
      oo.seg_reg_max = seg_cand[j][0] == 64 ? 0 : seg_cand[j][0];
      oo.seg_count_min = seg_cand[j][1] == 512 ? 0 : seg_cand[j][1];
      apply_seg_classes(c, oo);
    }
    // 6. nodal neighbour-count classes over every snapshot (the batch's nodal kernels)
    {
      const NodalBounds nc[] = {{16, 32, 128}, {16, 32, 64}, {16, 32, 96}, {-1, 32, 128},
                                {16, -1, 128}, {8, 32, 128}};
      std::vector<double> cost(6, 0.0);
      for (int j = 0; j < 6; ++j) {
        build_batch_lists(c, g, oo.num_batches, nc[j]);
        for (size_t k = 0; k < snaps.size(); ++k) cost[j] += wt[k] * t_batch(snaps[k], false);
      }
      const NodalBounds b = nc[pick(cost, 0)];
      oo.nodal_max16 = b.max16 == 16 ? 0 : b.max16;
      oo.nodal_max32 = b.max32 == 32 ? 0 : b.max32;
      oo.nodal_maxwave = b.maxwave == 128 ? 0 : b.maxwave;
    }
    HIPX(hipStreamSynchronize(st));
    o->tile_max_B = tile;
    o->dense_max_B = dense;
    o->seg_reg_max = oo.seg_reg_max;
    o->seg_count_min = oo.seg_count_min;
    o->nodal_max16 = oo.nodal_max16;
    o->nodal_max32 = oo.nodal_max32;
    o->nodal_maxwave = oo.nodal_maxwave;
  });
}

int gsap_get_trace(const gsap_ctx *c, gsap_trace_rec *out, int64_t cap, int64_t *n) {
  if (!c || !n) return GSAP_ERR_INVALID;
  *n = (int64_t)c->trace.size();
  if (out)
    for (int64_t i = 0; i < *n && i < cap; ++i) out[i] = c->trace[i];
  return GSAP_OK;
}

// the current block matrix (B blocks) into caller arrays
static void copy_blockmodel(gsap_ctx *c, int32_t B, int64_t *rowptr, int32_t *col, int64_t *val,
                            int64_t *colptr, int32_t *row, int64_t *cval, int64_t *dout,
                            int64_t *din, int64_t *nnz_out) {
    const int32_t nnz = read_nnz(c);
    if (nnz < 0 || nnz > c->E) throw Err(GSAP_ERR_DEVICE, "block matrix entry count out of range");
    *nnz_out = nnz;
    std::vector<int32_t> rp(B + 1), cp(B + 1), v(nnz), cv(nnz), dO(B), dI(B);
    hipStream_t st = c->st;
    HIPX(hipMemcpyAsync(rp.data(), c->bm.rowptr, 4 * (size_t)(B + 1), hipMemcpyDeviceToHost, st));
    HIPX(hipMemcpyAsync(cp.data(), c->bm.colptr, 4 * (size_t)(B + 1), hipMemcpyDeviceToHost, st));
    if (nnz > 0) {
      HIPX(hipMemcpyAsync(col, c->bm.col, 4 * (size_t)nnz, hipMemcpyDeviceToHost, st));
      HIPX(hipMemcpyAsync(row, c->bm.row, 4 * (size_t)nnz, hipMemcpyDeviceToHost, st));
      HIPX(hipMemcpyAsync(v.data(), c->bm.val, 4 * (size_t)nnz, hipMemcpyDeviceToHost, st));
      HIPX(hipMemcpyAsync(cv.data(), c->bm.cval, 4 * (size_t)nnz, hipMemcpyDeviceToHost, st));
    }
    HIPX(hipMemcpyAsync(dO.data(), c->bm.dout, 4 * (size_t)B, hipMemcpyDeviceToHost, st));
    HIPX(hipMemcpyAsync(dI.data(), c->bm.din, 4 * (size_t)B, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    for (int32_t x = 0; x <= B; ++x) { rowptr[x] = rp[x]; colptr[x] = cp[x]; }
    for (int32_t k = 0; k < nnz; ++k) { val[k] = v[k]; cval[k] = cv[k]; }
    for (int32_t x = 0; x < B; ++x) { dout[x] = dO[x]; din[x] = dI[x]; }
}

int gsap_blockmodel_i32(gsap_ctx *c, const gsap_graph_i32 *g, const int32_t *labels, int32_t B,
                        int64_t *rowptr, int32_t *col, int64_t *val, int64_t *colptr, int32_t *row,
                        int64_t *cval, int64_t *dout, int64_t *din, int64_t *nnz_out,
                        const gsap_options *o) {
  return guarded(c, [&] {
    if (!labels || !nnz_out) throw Err(GSAP_ERR_INVALID, "NULL argument");
    setup_unit(c, g, o, labels, B);
    copy_blockmodel(c, B, rowptr, col, val, colptr, row, cval, dout, din, nnz_out);
  });
}

int gsap_nodal_batch_bm_i32(gsap_ctx *c, const gsap_graph_i32 *g, const gsap_options *o,
                            int32_t *labels, int32_t B, int32_t b, uint32_t q, int32_t keep,
                            int64_t *rowptr, int32_t *col, int64_t *val, int64_t *colptr,
                            int32_t *row, int64_t *cval, int64_t *dout, int64_t *din,
                            int64_t *nnz_out) {
  return guarded(c, [&] {
    if (!labels || !nnz_out) throw Err(GSAP_ERR_INVALID, "NULL argument");
    gsap_options oo = o ? *o : defaults();
    if (b < 0 || b >= oo.num_batches) throw Err(GSAP_ERR_INVALID, "batch index out of range");
    if (!keep) {
      setup_unit(c, g, &oo, labels, B);
    } else {
      // the context's labels and block matrix as the previous call left them
      if (c->N != g->N || c->E != g->E || c->bm.B != B || c->lists_nb != oo.num_batches)
        throw Err(GSAP_ERR_INVALID, "keep = 1 needs the previous call's graph, B and batches");
      check_options(oo);
      set_device(c);
    }
    int32_t ar = avg_row(read_nnz(c), B);
    HIPX(hipMemcpyAsync(c->ctr.p, &q, 4, hipMemcpyHostToDevice, c->st));
    nodal_batch(c, oo, b, PhaseRef{c->ctr.as<uint32_t>(), 0}, ar, c->res.as<double>());
    rebuild_after_batch(c, B, oo, b);
    HIPX(hipMemcpyAsync(labels, c->lab.p, 4 * (size_t)c->N, hipMemcpyDeviceToHost, c->st));
    copy_blockmodel(c, B, rowptr, col, val, colptr, row, cval, dout, din, nnz_out);
  });
}

int gsap_merge_phase_i32(gsap_ctx *c, const gsap_graph_i32 *g, const gsap_options *o,
                         int32_t *labels, int32_t B, int32_t B_to_merge, uint32_t m, int32_t *prop,
                         double *dS, int32_t *best, double *dS_best, int32_t *B_new) {
  return guarded(c, [&] {
    if (!labels || !B_new) throw Err(GSAP_ERR_INVALID, "NULL argument");
    gsap_options oo = o ? *o : defaults();
    setup_unit(c, g, &oo, labels, B);
    std::vector<int32_t> hp, hb;
    std::vector<double> hd, hdb;
    *B_new = merge_phase(c, oo, B_to_merge, m, &hp, &hd, &hb, &hdb);
    HIPX(hipMemcpyAsync(labels, c->lab.p, 4 * (size_t)c->N, hipMemcpyDeviceToHost, c->st));
    HIPX(hipStreamSynchronize(c->st));
    if (prop) std::copy(hp.begin(), hp.end(), prop);
    if (dS) std::copy(hd.begin(), hd.end(), dS);
    if (best) std::copy(hb.begin(), hb.end(), best);
    if (dS_best) std::copy(hdb.begin(), hdb.end(), dS_best);
  });
}

int gsap_nodal_batch_i32(gsap_ctx *c, const gsap_graph_i32 *g, const gsap_options *o,
                         int32_t *labels, int32_t B, int32_t b, uint32_t q, int32_t *prop,
                         double *dS, double *dE) {
  return guarded(c, [&] {
    if (!labels || !dE) throw Err(GSAP_ERR_INVALID, "NULL argument");
    gsap_options oo = o ? *o : defaults();
    if (b < 0 || b >= oo.num_batches) throw Err(GSAP_ERR_INVALID, "batch index out of range");
    setup_unit(c, g, &oo, labels, B);
    const int64_t BS = c->N / oo.num_batches;
    int32_t ar = avg_row(read_nnz(c), B);
    double *res = c->res.as<double>();
    HIPX(hipMemcpyAsync(c->ctr.p, &q, 4, hipMemcpyHostToDevice, c->st));
    nodal_batch(c, oo, b, PhaseRef{c->ctr.as<uint32_t>(), 0}, ar, res);
    hipStream_t st = c->st;
    HIPX(hipMemcpyAsync(dE, res, 8, hipMemcpyDeviceToHost, st));
    if (prop && BS > 0)
      HIPX(hipMemcpyAsync(prop, c->prop.p, 4 * (size_t)BS, hipMemcpyDeviceToHost, st));
    if (dS && BS > 0) HIPX(hipMemcpyAsync(dS, c->dS.p, 8 * (size_t)BS, hipMemcpyDeviceToHost, st));
    HIPX(hipMemcpyAsync(labels, c->lab.p, 4 * (size_t)c->N, hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
  });
}

int gsap_mdl_i32(gsap_ctx *c, const gsap_graph_i32 *g, const int32_t *labels, int32_t B,
                 double *S, const gsap_options *o) {
  return guarded(c, [&] {
    if (!labels || !S) throw Err(GSAP_ERR_INVALID, "NULL argument");
    setup_unit(c, g, o, labels, B);
    *S = mdl_now(c, o ? o->model : GSAP_MODEL_REFERENCE);
  });
}

int gsap_mdl_noalloc_i32(gsap_ctx *c, const int32_t *d_out_ptr, const int32_t *d_out_adj,
                         const int32_t *d_out_w, const int32_t *d_labels, int64_t N, int64_t E,
                         int32_t B, int32_t model, double *d_S, void *stream) {
  return guarded(c, [&] {
    set_device(c);
    std::string msg;
    const int rc = noalloc_mdl(c->device, d_out_ptr, d_out_adj, d_out_w, d_labels, N, E, B,
                               model, d_S, stream ? (hipStream_t)stream : c->st, msg);
    if (rc != GSAP_OK) throw Err(rc, msg);
  });
}

}  // extern "C"
