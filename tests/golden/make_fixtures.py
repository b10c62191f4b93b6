#!/usr/bin/env python3
"""Generate the committed fixtures under tests/golden/.

    python tests/golden/make_fixtures.py

Inputs (seeded, see for-qaoa_amd/forqaoa/graphs.py):
  instances.json      the benchmark instances (edges, per-layer angles) and the
                      small instances used against the dense simulator.
Expected outputs, each from a method independent of the oracle:
  dense_small.npz     final states (logical qubit order) of the small
                      instances from tests/refmath.dense_qaoa (full-matrix
                      textbook gates, scipy expm).
  kn_closed_form.json phi(h) and <C> for K_N with unit weights and uniform
                      angles (Hamming-weight closed form, mpmath 40 digits),
                      including the recipe of the reference's own validation
                      script (N=30, P=1, gamma=beta=pi/4,
                      scripts/valiate_ans_scripts/validate_state.py:12-46).
  p1_analytic.json    single-layer <C> of the N=20 and N=30 benchmark graphs
                      from the closed-form expectation (tests/refmath.py).
Regression fixture produced by the oracle itself (pins nothing on its own; it
freezes the oracle's N=20 output so the GPU can be compared without re-running
it, and so any later change to the oracle shows up):
  oracle_n20.json     N=20, P=5, C=12: <C>, norm, top-8, qubit map and 512
                      sampled amplitudes (physical order).
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "for-qaoa_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from forqaoa import graphs as G  # noqa: E402
import refmath as R  # noqa: E402

# small instances for the dense simulator: (name, N, C, P, generator)
SMALL = [
    ("s4_gnp", 4, 2, 3, ("gnp", 0.8, "uniform")),
    ("s5_gnp_c5", 5, 5, 2, ("gnp", 0.6, "uniform")),
    ("s6_reg3", 6, 4, 2, ("reg", 3, "uniform")),
    ("s7_gnp_int", 7, 3, 3, ("gnp", 0.5, "int16")),
    ("s8_gnp", 8, 3, 3, ("gnp", 0.6, "uniform")),
    ("s9_gnp_c4", 9, 4, 2, ("gnp", 0.4, "uniform")),
    ("s10_gnp", 10, 4, 2, ("gnp", 0.5, "uniform")),
    ("s10_reg4_c10", 10, 10, 2, ("reg", 4, "uniform")),
]

BENCH = ["n20_gnp05", "n30_reg3", "n30_gnp05", "n34_gnp03_int", "n32_gnp03_int", "n26_reg4"]


def small_instance(name, N, C, P, gen, seed):
    kind, a, wk = gen
    edges = G.gnp(N, a, seed, wk) if kind == "gnp" else G.regular(a, N, seed, wk)
    g, b = G.angles(P, seed)
    return dict(name=name, N=N, C=C, P=P, seed=seed, edges=edges, gammas=list(g), betas=list(b))


def main():
    inst = {}
    for i, (name, N, C, P, gen) in enumerate(SMALL):
        inst[name] = small_instance(name, N, C, P, gen, G.BASE_SEED + 100 + i)
    for name in BENCH:
        w = G.workload(name)
        inst[name] = dict(name=name, N=w["N"], C=w["C"], P=w["P"], seed=w["seed"], edges=w["edges"],
                          gammas=list(map(float, w["gammas"])), betas=list(map(float, w["betas"])))
    with open(os.path.join(HERE, "instances.json"), "w") as f:
        json.dump(inst, f, indent=1)

    dense = {}
    for name, *_ in SMALL:
        d = inst[name]
        dense[name] = R.dense_qaoa(d["N"], d["edges"], d["gammas"], d["betas"])
    np.savez_compressed(os.path.join(HERE, "dense_small.npz"), **dense)

    kn = []
    cases = [
        (12, [math.pi / 4] * 3, [math.pi / 4] * 3, "pi/4 x3"),
        (16, None, None, "random P=2"),
        (20, [math.pi / 4] * 5, [math.pi / 4] * 5, "pi/4 x5"),
        (30, [math.pi / 4], [math.pi / 4], "validate_state.py recipe: N=30 P=1 pi/4"),
        (30, None, None, "random P=5"),
        (34, None, None, "random P=2"),
    ]
    for N, gs, bs, label in cases:
        if gs is None:
            P = 5 if N == 30 else 2
            ga, be = G.angles(P, G.BASE_SEED + 500 + N)
            gs, bs = list(map(float, ga)), list(map(float, be))
        phi = R.kn_phi(N, gs, bs)
        ex = R.kn_expect_cut(N, phi)
        kn.append(dict(N=N, gammas=gs, betas=bs, label=label,
                       phi_re=[float(p.real) for p in phi], phi_im=[float(p.imag) for p in phi],
                       expect_cut=float(ex)))
    with open(os.path.join(HERE, "kn_closed_form.json"), "w") as f:
        json.dump(kn, f, indent=1)

    p1 = {}
    for name in ["n20_gnp05", "n30_reg3", "n30_gnp05"]:
        d = inst[name]
        p1[name] = dict(gamma=d["gammas"][0], beta=d["betas"][0],
                        expect_cut=R.p1_expect_cut(d["N"], d["edges"], d["gammas"][0], d["betas"][0]))
    with open(os.path.join(HERE, "p1_analytic.json"), "w") as f:
        json.dump(p1, f, indent=1)

    import pyoracle as O

    d = inst["n20_gnp05"]
    sv, qmap, W = O.run_weighted(d["N"], d["C"], d["edges"], d["gammas"], d["betas"])
    ex = O.expect_cut(sv, W)
    n2 = O.norm2(sv)
    idx, prob = O.topk(sv, 8)
    lidx = [int(sum(((int(b) >> p) & 1) << int(qmap[p]) for p in range(d["N"]))) for b in idx]
    rng = np.random.default_rng(7)
    samp = sorted(set(int(x) for x in rng.integers(0, 1 << d["N"], size=512)))
    with open(os.path.join(HERE, "oracle_n20.json"), "w") as f:
        json.dump(dict(name="n20_gnp05", expect_cut=ex, norm2=n2, qmap=[int(q) for q in qmap],
                       top8_phys=[int(i) for i in idx], top8_logical=lidx, top8_prob=[float(p) for p in prob],
                       sample_idx=samp, sample_re=[float(sv[i].real) for i in samp],
                       sample_im=[float(sv[i].imag) for i in samp]), f, indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
