// Drop-in replacement of the reference's partition() (GSAP/partition.hpp:852-954).
//
//   #include "gsap/graph.hpp"
//   #include "gsap/partition.hpp"
//   partition(g);                 // g.partitions <- block ids, like the reference
//
// The call packs g.csr_out / g.csr_in (appending the end pointer the reference's Csr
// omits) into a gsap_graph_i32 and runs the HIP implementation through the C ABI of
// gsap.h on the calling thread's context (device 0 unless gsap::set_device() chose
// another).  Errors raise std::runtime_error; the C ABI itself never throws.
// Link with libgsap_amd.so.
#pragma once

#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "gsap.h"
#include "graph.hpp"

namespace gsap {

struct Context {
  gsap_ctx *h = nullptr;
  int device = -1;
  ~Context() {
    if (h) gsap_destroy(h);
  }
};

inline int &thread_device() {
  thread_local int d = 0;
  return d;
}

inline void set_device(int d) { thread_device() = d; }

inline gsap_ctx *thread_context() {
  thread_local Context c;
  if (!c.h || c.device != thread_device()) {
    if (c.h) gsap_destroy(c.h);
    c.h = nullptr;
    int rc = gsap_create(&c.h, thread_device());
    if (rc != GSAP_OK)
      throw std::runtime_error("gsap_create(device " + std::to_string(thread_device()) +
                               ") failed with status " + std::to_string(rc));
    c.device = thread_device();
  }
  return c.h;
}

inline gsap_options default_options() {
  gsap_options o;
  gsap_options_default(&o);
  return o;
}

template <typename T, typename V, typename W>
void partition(Graph<T, V, W> &g, const gsap_options &opts, gsap_report *report = nullptr) {
  static_assert(std::is_integral<T>::value && std::is_integral<V>::value &&
                    std::is_integral<W>::value,
                "Graph<T,V,W> must use integral types");
  static_assert(sizeof(V) <= 4 && sizeof(T) <= 8, "vertex ids must fit 32 bits");
  const int64_t N = g.N;
  const int64_t E = (int64_t)g.csr_out.adj_node.size();
  if (N <= 0) throw std::runtime_error("partition: empty graph");
  if ((int64_t)g.csr_out.adj_ptr.size() != N || (int64_t)g.csr_in.adj_ptr.size() != N ||
      (int64_t)g.csr_in.adj_node.size() != E || (int64_t)g.csr_out.adj_wgt.size() != E ||
      (int64_t)g.csr_in.adj_wgt.size() != E)
    throw std::runtime_error("partition: csr_out / csr_in sizes do not match N and E");
  std::vector<int32_t> op(N + 1), ip(N + 1), oa(E), ia(E), ow(E), iw(E);
  for (int64_t v = 0; v < N; ++v) {
    op[v] = (int32_t)g.csr_out.adj_ptr[v];
    ip[v] = (int32_t)g.csr_in.adj_ptr[v];
  }
  op[N] = (int32_t)E;
  ip[N] = (int32_t)E;
  for (int64_t e = 0; e < E; ++e) {
    oa[e] = (int32_t)g.csr_out.adj_node[e];
    ia[e] = (int32_t)g.csr_in.adj_node[e];
    if ((uint64_t)g.csr_out.adj_wgt[e] > (uint64_t)std::numeric_limits<int32_t>::max() ||
        (uint64_t)g.csr_in.adj_wgt[e] > (uint64_t)std::numeric_limits<int32_t>::max())
      throw std::runtime_error("partition: edge weight does not fit int32");
    ow[e] = (int32_t)g.csr_out.adj_wgt[e];
    iw[e] = (int32_t)g.csr_in.adj_wgt[e];
  }
  gsap_graph_i32 cg{N, E, op.data(), oa.data(), ow.data(), ip.data(), ia.data(), iw.data()};
  std::vector<int32_t> out(N);
  gsap_ctx *ctx = thread_context();
  int rc = gsap_partition_i32(ctx, &cg, &opts, out.data(), report);
  if (rc != GSAP_OK)
    throw std::runtime_error(std::string("partition: ") + gsap_last_error(ctx));
  g.partitions.assign(out.begin(), out.end());
}

}  // namespace gsap

// Same spelling as the reference: template<T,V,W> void partition(Graph<T,V,W>&).
template <typename T, typename V, typename W>
void partition(Graph<T, V, W> &g) {
  gsap::partition(g, gsap::default_options());
}
