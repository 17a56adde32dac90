"""The benchmark workloads: seeded DC-SBM graphs of the five configurations.

  c1  N = 1,000, 11 near-equal blocks, intra/inter ratio 4, mean out-degree 20, seed 1
  c2  N = 20,000, 33 Dirichlet(1) blocks, ratio 1.5, mean 20, seed 2
  c3  N = 200,000, 71 Dirichlet(1) blocks, ratio 1.5, mean 22.5 (E = 4.5M), seed 3
  c4  256 graphs, N log-uniform in [500, 50,000], K = round(N^(1/3)), Dirichlet(1) sizes,
      ratio 4 (even g) / 1.5 (odd g), mean 20, seed 4 + g
  c5  N = 2,000,000, 150 Dirichlet(0.5) blocks, ratio 2, mean 22.5 (E = 45M), seed 5

Every graph is a pure function of these parameters (the generator is Philox-seeded).
"""
from __future__ import annotations

import numpy as np

from . import generate_dcsbm

CONFIGS = {
    "c1": dict(N=1000, K=11, alpha=0.0, rho=4.0, mean_deg=20.0, seed=1),
    "c2": dict(N=20000, K=33, alpha=1.0, rho=1.5, mean_deg=20.0, seed=2),
    "c3": dict(N=200000, K=71, alpha=1.0, rho=1.5, mean_deg=22.5, seed=3),
    "c5": dict(N=2000000, K=150, alpha=0.5, rho=2.0, mean_deg=22.5, seed=5),
}


def graph(name: str):
    """One of the single-graph configurations c1, c2, c3, c5."""
    c = CONFIGS[name]
    return generate_dcsbm(c["N"], c["K"], c["alpha"], c["rho"], c["mean_deg"], seed=c["seed"])


def c4_params(n: int = 256):
    """(N, K, rho, seed) of the batched configuration's graphs."""
    rng = np.random.default_rng(4)
    Ns = np.exp(rng.uniform(np.log(500), np.log(50000), n)).astype(int)
    return [(int(N), max(2, int(round(N ** (1 / 3)))), 4.0 if g % 2 == 0 else 1.5, 4 + g)
            for g, N in enumerate(Ns)]


def c4_graph(g: int, n: int = 256):
    N, K, rho, seed = c4_params(n)[g]
    return generate_dcsbm(N, K, 1.0, rho, 20.0, seed=seed)


def c4_graphs(n: int = 256):
    """The batched configuration: 256 independent graphs."""
    return [c4_graph(g, n) for g in range(n)]
