"""Seeded synthetic Max-Cut instances (graphs and per-layer angles).

The reference hard-codes the complete graph K_N with unit weights and one
angle gamma = beta = pi/4 for every layer (src/Weighted.cpp:34-35, 86-88).
These generators supply the weighted instances of the benchmark workloads:

* G(n, p): for every pair i < j in lexicographic order draw rng.random() and
  keep the edge when it is < p; then draw one weight per kept edge, in the
  same order.
* k-regular: networkx.random_regular_graph(k, n, seed), edges sorted, then one
  weight per edge from the numpy generator.
* weights "unit" (1.0), "uniform" (1.0 - rng.random(), i.e. in (0, 1]) or
  "int16" (rng.integers(1, 17), i.e. in {1..16}).
* angles gamma_k, beta_k = rng.uniform(0, pi), drawn layer by layer from a
  generator seeded with seed + 1000.

Edges are (u, v, w) tuples with u < v.
"""
from __future__ import annotations

import math

import numpy as np

BASE_SEED = 20251202


def _weights(rng, m, kind):
    if kind == "unit":
        return [1.0] * m
    if kind == "uniform":
        return [float(1.0 - rng.random()) for _ in range(m)]
    if kind == "int16":
        return [float(rng.integers(1, 17)) for _ in range(m)]
    raise ValueError(f"unknown weight kind {kind!r}")


def gnp(n, p, seed, weights="uniform"):
    rng = np.random.default_rng(seed)
    pairs = []
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < p:
                pairs.append((i, j))
    ws = _weights(rng, len(pairs), weights)
    return [(i, j, w) for (i, j), w in zip(pairs, ws)]


def regular(k, n, seed, weights="uniform"):
    import networkx as nx

    g = nx.random_regular_graph(k, n, seed=seed)
    pairs = sorted((min(u, v), max(u, v)) for u, v in g.edges())
    rng = np.random.default_rng(seed)
    ws = _weights(rng, len(pairs), weights)
    return [(i, j, w) for (i, j), w in zip(pairs, ws)]


def complete(n, w=1.0):
    """K_n with constant weight: the reference's own instance."""
    return [(i, j, float(w)) for i in range(n) for j in range(i + 1, n)]


def angles(P, seed):
    rng = np.random.default_rng(seed + 1000)
    g = np.zeros(P)
    b = np.zeros(P)
    for k in range(P):
        g[k] = rng.uniform(0, math.pi)
        b[k] = rng.uniform(0, math.pi)
    return g, b


def total_weight(edges):
    return float(sum(w for _, _, w in edges))


def workload(name):
    """Named instances of the benchmark configurations.

    Returns dict(N, P, C, edges, gammas, betas, seed, name, desc)."""
    table = {
        # N=20, P=5, C=12, G(20, 0.5), weights in (0, 1]
        "n20_gnp05": dict(N=20, P=5, C=12, seed=BASE_SEED + 1, gen=("gnp", 0.5, "uniform")),
        # N=30, P=5, 3-regular and G(30, 0.5), weights in (0, 1]
        "n30_reg3": dict(N=30, P=5, C=13, seed=BASE_SEED + 2, gen=("reg", 3, "uniform")),
        "n30_gnp05": dict(N=30, P=5, C=13, seed=BASE_SEED + 2, gen=("gnp", 0.5, "uniform")),
        # N=34, P=10, G(34, 0.3) integer weights; N=32 on the same generator
        "n34_gnp03_int": dict(N=34, P=10, C=14, seed=BASE_SEED + 3, gen=("gnp", 0.3, "int16")),
        "n32_gnp03_int": dict(N=32, P=10, C=13, seed=BASE_SEED + 3, gen=("gnp", 0.3, "int16")),
        # N=26, P=8, 4-regular, weights in (0, 1]
        "n26_reg4": dict(N=26, P=8, C=14, seed=BASE_SEED + 4, gen=("reg", 4, "uniform")),
    }
    if name not in table:
        raise KeyError(name)
    spec = dict(table[name])
    kind, a, wk = spec.pop("gen")
    if kind == "gnp":
        edges = gnp(spec["N"], a, spec["seed"], wk)
    else:
        edges = regular(a, spec["N"], spec["seed"], wk)
    g, b = angles(spec["P"], spec["seed"])
    wdesc = {"uniform": "weights U(0,1]", "int16": "integer weights 1..16", "unit": "unit weights"}[wk]
    gdesc = f"G({spec['N']},{a})" if kind == "gnp" else f"{a}-regular"
    spec.update(edges=edges, gammas=g, betas=b, name=name,
                desc=f"N={spec['N']} P={spec['P']} C={spec['C']} {gdesc} {wdesc}")
    return spec
