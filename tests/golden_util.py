"""Loaders for the committed fixtures in tests/golden/ (test infrastructure)."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def instances():
    with open(os.path.join(GOLDEN, "instances.json")) as f:
        d = json.load(f)
    for v in d.values():
        v["edges"] = [(int(u), int(w_), float(x)) for u, w_, x in v["edges"]]
    return d


def dense_small():
    z = np.load(os.path.join(GOLDEN, "dense_small.npz"))
    return {k: z[k] for k in z.files}


def kn_cases():
    with open(os.path.join(GOLDEN, "kn_closed_form.json")) as f:
        d = json.load(f)
    for c in d:
        c["phi"] = np.array(c["phi_re"]) + 1j * np.array(c["phi_im"])
    return d


def p1_analytic():
    with open(os.path.join(GOLDEN, "p1_analytic.json")) as f:
        return json.load(f)


def oracle_n20():
    with open(os.path.join(GOLDEN, "oracle_n20.json")) as f:
        return json.load(f)


def popcount(idx):
    idx = np.asarray(idx, dtype=np.uint64)
    c = np.zeros(idx.shape, dtype=np.int64)
    x = idx.copy()
    while np.any(x):
        c += (x & np.uint64(1)).astype(np.int64)
        x >>= np.uint64(1)
    return c


def logical_index(b, qmap):
    return sum(((int(b) >> p) & 1) << int(qmap[p]) for p in range(len(qmap)))
