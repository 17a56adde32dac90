// Graph types of the drop-in C++ API.  Field names and shapes match the reference's
// GSAP/graph.hpp:8-45 (Edge, Csr, Graph) so code written against it compiles unchanged:
//   Csr::adj_ptr has N entries (row starts, no end sentinel), adj_node / adj_wgt have E.
#pragma once

#include <utility>
#include <vector>

template <typename V, typename W>
struct Edge {
  V from;
  V to;
  W wgt;
  Edge(V f, V t, W w) : from(f), to(t), wgt(w) {}
};

template <typename T, typename V, typename W>
struct Csr {
  std::vector<T> adj_ptr;   // N row starts
  std::vector<V> adj_node;  // E neighbour ids, 0-based
  std::vector<W> adj_wgt;   // E weights
  std::vector<W> deg;       // unused by partition()

  void reset(int B) {
    adj_ptr.clear();
    adj_node.clear();
    adj_wgt.clear();
    deg.assign(B, W());
  }
};

template <typename T, typename V, typename W>
struct Graph {
  std::vector<V> partitions;      // output of partition(): N block ids
  std::vector<V> truePartitions;  // optional planted labels (not read by partition())
  int N = 0;
  int E = 0;
  std::vector<Edge<V, W>> edges;
  std::vector<std::vector<std::pair<V, W>>> out_neighbors;
  std::vector<std::vector<std::pair<V, W>>> in_neighbors;
  Csr<T, V, W> csr_out;
  Csr<T, V, W> csr_in;
  std::vector<W> node_deg_out;
  std::vector<W> node_deg_in;
};
