#!/usr/bin/env python3
"""Oracle fixtures for the full-size benchmark instances.

    python tests/golden/make_oracle_large.py [n30_gnp05] [n30_reg3] [n26_reg4]

Runs the CPU oracle (oracle/fq_oracle.c: the reference's layer loop,
src/Weighted.cpp:97-137, in complex128) once per instance and freezes what the
GPU runs are compared with, so the GPU tests need not run the oracle:

  oracle_n30_gnp05.npz, oracle_n30_reg3.npz
      N=30, P=5, C=13 (the instances of tests/golden/instances.json):
      expect_cut, norm2, the 16 most probable LOGICAL bitstrings and their
      probabilities (top16_idx, top16_prob), the final qubit map, and 8192
      sampled amplitudes at LOGICAL indices (sample_idx, sample_re, sample_im).
      N=30 complex128 holds a 17.2 GB state: about 20 GB of host memory.
  oracle_n26_reg4.npz
      N=26, P=8, C=14, 4-regular: <C> at the 16 angle sets drawn by
      numpy.random.default_rng(2026) (gamma then beta, P values each, uniform
      in [0, pi)), the sets the adjoint workload test uses.

The oracle is pinned independently (tests/test_oracle_pins.py); these files
spare the GPU tests a 5-minute CPU run.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "for-qaoa_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import golden_util as GU  # noqa: E402
import pyoracle as O  # noqa: E402

N26_ANGLE_SEED = 2026
N26_SETS = 16
SAMPLE_SEED = 30
SAMPLES = 8192


def phys_of_logical(idx, qmap):
    """Physical index of each logical index: bit qmap[p] of the logical index
    is bit p of the physical one."""
    idx = np.asarray(idx, dtype=np.uint64)
    out = np.zeros_like(idx)
    for p, q in enumerate(qmap):
        out |= ((idx >> np.uint64(q)) & np.uint64(1)) << np.uint64(p)
    return out


def logical_of_phys(idx, qmap):
    idx = np.asarray(idx, dtype=np.uint64)
    out = np.zeros_like(idx)
    for p, q in enumerate(qmap):
        out |= ((idx >> np.uint64(p)) & np.uint64(1)) << np.uint64(q)
    return out


def n30(name):
    d = GU.instances()[name]
    N, C = d["N"], d["C"]
    t0 = time.time()
    sv, qmap, W = O.run_weighted(N, C, d["edges"], d["gammas"], d["betas"])
    t1 = time.time()
    ex = O.expect_cut(sv, W)
    n2 = O.norm2(sv)
    pidx, prob = O.topk(sv, 16)
    rng = np.random.default_rng(SAMPLE_SEED)
    lidx = np.unique(rng.integers(0, 1 << N, size=SAMPLES, dtype=np.uint64))
    amps = sv[phys_of_logical(lidx, qmap).astype(np.int64)]
    out = os.path.join(HERE, f"oracle_{name}.npz")
    np.savez_compressed(out, expect_cut=np.float64(ex), norm2=np.float64(n2), qmap=qmap.astype(np.int32),
                        top16_idx=logical_of_phys(pidx, qmap), top16_prob=prob,
                        sample_idx=lidx, sample_re=amps.real.copy(), sample_im=amps.imag.copy())
    print(f"{name}: oracle loop {t1 - t0:.1f} s on {O.num_threads()} threads, <C> = {ex!r}, "
          f"norm2 - 1 = {n2 - 1:.2e} -> {out}", flush=True)


def n26():
    d = GU.instances()["n26_reg4"]
    N, C, P = d["N"], d["C"], d["P"]
    rng = np.random.default_rng(N26_ANGLE_SEED)
    gs, bs, ex = [], [], []
    for s in range(N26_SETS):
        g = rng.uniform(0, np.pi, P)
        b = rng.uniform(0, np.pi, P)
        sv, qmap, W = O.run_weighted(N, C, d["edges"], g, b)
        ex.append(O.expect_cut(sv, W))
        gs.append(g)
        bs.append(b)
        print(f"n26_reg4 set {s}: <C> = {ex[-1]!r}", flush=True)
    out = os.path.join(HERE, "oracle_n26_reg4.npz")
    np.savez_compressed(out, gammas=np.array(gs), betas=np.array(bs), expect_cut=np.array(ex))
    print("->", out)


def main(argv):
    todo = argv or ["n26_reg4", "n30_gnp05", "n30_reg3"]
    for name in todo:
        if name == "n26_reg4":
            n26()
        elif name in ("n30_gnp05", "n30_reg3"):
            n30(name)
        else:
            raise SystemExit(f"unknown instance {name}")


if __name__ == "__main__":
    main(sys.argv[1:])
