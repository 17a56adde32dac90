// Checker driver (test infrastructure only): runs the reference's own evaluation code,
// GSAP/evaluate/evaluate.hpp and km.hpp compiled from the reference checkout, on two
// labellings and prints its metrics, then the total of the best one-to-one matching of
// truth and given blocks found by the reference's Kuhn-Munkres solver.  That solver
// maximises total weight (km.hpp:70-74 starts the row labels at the row maxima), so it is
// given the contingency table itself; evaluate.hpp:62-71 hands it the negated table, which
// makes the matching it uses the worst one (its accuracy print is commented out).
// Input file: N, then N truth labels (0-based, contiguous), then N given labels.
// Built only into oracle/_ref/ by oracle/Makefile; nothing in the product uses it.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <vector>

#include "evaluate.hpp"

int main(int argc, char **argv) {
  if (argc != 2) return 2;
  std::ifstream f(argv[1]);
  long n = 0;
  f >> n;
  std::vector<int> t(n), g(n);
  for (auto &x : t) f >> x;
  for (auto &x : g) f >> x;
  if (!f) return 3;
  bf::evaluate<int>(t, g);
  const int Bt = *std::max_element(t.begin(), t.end()) + 1;
  const int Bg = *std::max_element(g.begin(), g.end()) + 1;
  const bool tr = Bt > Bg;  // the solver expects at least as many columns as rows
  const int R = tr ? Bg : Bt, C = tr ? Bt : Bg;
  std::vector<std::vector<int>> table(R, std::vector<int>(C, 0));
  for (long i = 0; i < n; ++i) {
    if (tr) table[g[i]][t[i]]++;
    else table[t[i]][g[i]]++;
  }
  bf::KuhnMunkresSolver<int> km;
  km.solve(table);
  const auto &m = km.match1();
  long total = 0;
  for (int i = 0; i < R; ++i) total += table[i][m[i]];
  std::printf("Matched total: %ld\n", total);
  return 0;
}
