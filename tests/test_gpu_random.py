"""Seeded random configurations against the oracle: qubit counts across the
tile-width boundaries, chunk sizes, graph densities and weight kinds, depths,
every schedule and both amplitude types (60 cases, fixed seed).  Each case
prints its parameters before running."""
import numpy as np
import pytest

import pyoracle as O
from forqaoa import Graph, State
from forqaoa import graphs as G

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


def _cases(n=60, seed=20251202):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        N = int(rng.integers(3, 25))
        dtype = ["c128", "c64"][i % 2]
        cmax = 13 if dtype == "c128" else 14
        C = int(rng.integers(1, min(N, cmax) + 1))
        schedule = ["reference", "fused", "z2"][int(rng.integers(0, 3))]
        P = int(rng.integers(1, 5))
        p = float(rng.uniform(0.2, 0.9))
        kind = ["uniform", "int16", "unit"][int(rng.integers(0, 3))]
        out.append((i, N, C, dtype, schedule, P, p, kind))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: f"{c[0]}-N{c[1]}-C{c[2]}-{c[3]}-{c[4]}-P{c[5]}")
def test_random_configuration_vs_oracle(case):
    i, N, C, dtype, schedule, P, p, kind = case
    print(f"case {i}: N={N} C={C} {dtype} {schedule} P={P} G(N,{p:.2f}) {kind}", flush=True)
    edges = G.gnp(N, p, 1000 + i, kind)
    g_, b_ = G.angles(P, 2000 + i)
    sv, qmap, W = O.run_weighted(N, C, edges, g_, b_)
    st = State(N, C=C, dtype=dtype)
    g = Graph(N, edges)
    try:
        st.run(g, g_, b_, schedule)
        got = st.copy_out(logical=True)
        ref = O.to_logical(sv, qmap)
        err = float(np.linalg.norm(got.astype(np.complex128) - ref) / np.linalg.norm(ref))
        assert err < TOL[dtype], err
        ex, n2 = st.expect_cut(g)
        # the reference schedule relabels g in place, so g matches the final layout
        assert ex == pytest.approx(O.expect_cut(sv, W), rel=TOL[dtype], abs=TOL[dtype])
        assert n2 == pytest.approx(1.0, abs=TOL[dtype] * 10)
    finally:
        st.destroy()
        g.destroy()
