// Driver in the shape of the reference's GSAP.cu:7-34, built against the drop-in headers:
// read a GraphChallenge TSV into Graph<int,int,unsigned>, time partition(g), print results.
#include <chrono>
#include <cstdio>
#include <vector>

#include "gsap/graph.hpp"
#include "gsap/partition.hpp"

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s graph.tsv\n", argv[0]);
    return 2;
  }
  int64_t N = 0, E = 0;
  if (gsap_read_tsv_i32(argv[1], &N, &E, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr))
    return 3;
  std::vector<int32_t> op(N + 1), ip(N + 1), oa(E), ia(E), ow(E), iw(E);
  if (gsap_read_tsv_i32(argv[1], &N, &E, op.data(), oa.data(), ow.data(), ip.data(), ia.data(),
                        iw.data()))
    return 3;
  Graph<int, int, unsigned> g;
  g.N = (int)N;
  g.E = (int)E;
  g.csr_out.adj_ptr.assign(op.begin(), op.end() - 1);  // N row starts, as load.hpp:89-111
  g.csr_in.adj_ptr.assign(ip.begin(), ip.end() - 1);
  g.csr_out.adj_node.assign(oa.begin(), oa.end());
  g.csr_in.adj_node.assign(ia.begin(), ia.end());
  g.csr_out.adj_wgt.assign(ow.begin(), ow.end());
  g.csr_in.adj_wgt.assign(iw.begin(), iw.end());
  std::printf("Number of Vertices                : %d\nNumber of Edges                   : %d\n",
              g.N, g.E);
  try {
    auto t0 = std::chrono::steady_clock::now();
    partition(g);
    auto t1 = std::chrono::steady_clock::now();
    std::printf("Partitioning Runtime (ms)         : %lld\n",
                (long long)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count());
  } catch (const std::exception &e) {
    std::fprintf(stderr, "partition failed: %s\n", e.what());
    return 1;
  }
  uint64_t h = 1469598103934665603ull;
  int B = 0;
  for (int v : g.partitions) {
    for (int k = 0; k < 4; ++k) { h ^= (unsigned char)((unsigned)v >> (8 * k)); h *= 1099511628211ull; }
    if (v + 1 > B) B = v + 1;
  }
  std::printf("labels %016llx max_block %d\n", (unsigned long long)h, B);
  return 0;
}
