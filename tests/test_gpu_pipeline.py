"""Pipelined layers of the fused and half-state schedules.

The pass that ends layer p applies RX(beta_p), the cost phase of layer p+1 and
RX(beta_{p+1}) on its qubits, so a run of P layers with m passes per layer
launches 1 + P(m-1) tile passes instead of Pm (FQ_LAYER_PIPE=0 turns it
off).  Checked here:

* against the oracle's state vector (an independent layer-by-layer path) for
  qubit counts with 2 and 3 passes per layer, odd and even depths (layers end
  alternately on pass 0 and on the last pass, fused, or pass 1, half state),
  both schedules, both amplitude types;
* against the same state run in plain layer order (FQ_LAYER_PIPE=0);
* the number of tile-pass launches, from the library's own launch counters.
"""
import os

import numpy as np
import pytest

import pyoracle as O
from forqaoa import KERNEL_CLASSES, Graph, State, layer_passes
from forqaoa import graphs as G

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}
FUSED = KERNEL_CLASSES.index("fused")


@pytest.fixture
def pipe_env():
    old = os.environ.get("FQ_LAYER_PIPE")
    yield
    if old is None:
        os.environ.pop("FQ_LAYER_PIPE", None)
    else:
        os.environ["FQ_LAYER_PIPE"] = old


def _run(N, dtype, schedule, P, seed, pipe):
    edges = G.gnp(N, 0.5, 3000 + seed, ["uniform", "int16", "unit"][seed % 3])
    g_, b_ = G.angles(P, 4000 + seed)
    if pipe is None:
        os.environ.pop("FQ_LAYER_PIPE", None)
    else:
        os.environ["FQ_LAYER_PIPE"] = pipe
    st = State(N, C=min(N, 10), dtype=dtype)
    g = Graph(N, edges)
    try:
        st.timing(True)
        st.timing_reset()
        st.run(g, g_, b_, schedule)
        st.sync()
        launches = st.timing_get(FUSED)[1]
        st.timing(False)
        return st.copy_out(logical=True), launches, edges, g_, b_
    finally:
        st.destroy()
        g.destroy()


CASES = [
    # N, dtype, schedule, P (2 passes per layer: N 14..23 complex128 fused)
    (14, "c128", "fused", 1), (14, "c128", "fused", 2), (17, "c128", "fused", 3),
    (20, "c128", "fused", 4), (23, "c128", "fused", 5),
    # 3 passes per layer (fused from N=24 complex128; half state from N=24)
    (24, "c128", "fused", 3), (24, "c128", "z2", 2), (24, "c128", "z2", 3), (25, "c128", "z2", 4),
    (24, "c128", "fused", 4),
    # complex64 (N=24, P=4: the contiguous pass 0 ends layers 1 and 3 as a
    # 1024 x 16 boundary pass whose qubits 4 and 5 get RX by lane swaps)
    (15, "c64", "fused", 3), (24, "c64", "fused", 2), (25, "c64", "z2", 3), (25, "c64", "z2", 4),
    (24, "c64", "fused", 4),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"N{c[0]}-{c[1]}-{c[2]}-P{c[3]}")
def test_pipelined_layers_vs_oracle(case, pipe_env):
    N, dtype, schedule, P = case
    seed = N * 10 + P
    print(f"pipelined N={N} {dtype} {schedule} P={P}", flush=True)
    got, launches, edges, g_, b_ = _run(N, dtype, schedule, P, seed, "1")
    m = len(layer_passes(N, dtype, schedule))
    assert m >= (3 if schedule == "z2" else 2)
    assert launches == 1 + P * (m - 1)
    sv, qmap, W = O.run_weighted(N, 4, edges, g_, b_)
    ref = O.to_logical(sv, qmap)
    err = float(np.linalg.norm(got.astype(np.complex128) - ref) / np.linalg.norm(ref))
    print(f"  {launches} tile passes, relative L2 error vs oracle {err:.2e}", flush=True)
    assert err < TOL[dtype]


@pytest.mark.parametrize("N,dtype,schedule,P", [(16, "c128", "fused", 3), (24, "c128", "z2", 4),
                                                (24, "c128", "fused", 5), (20, "c64", "fused", 4)])
def test_pipelined_matches_layer_order(N, dtype, schedule, P, pipe_env):
    a, la, *_ = _run(N, dtype, schedule, P, 7, "1")
    b, lb, *_ = _run(N, dtype, schedule, P, 7, "0")
    m = len(layer_passes(N, dtype, schedule))
    assert (la, lb) == (1 + P * (m - 1), P * m)
    err = float(np.max(np.abs(a.astype(np.complex128) - b)) / np.max(np.abs(b)))
    print(f"N={N} {dtype} {schedule} P={P}: {la} vs {lb} passes, max diff / max |psi| {err:.2e}", flush=True)
    assert err < TOL[dtype]


def test_pipelined_by_default(pipe_env):
    for dtype in ("c128", "c64"):
        _, launches, *_ = _run(24, dtype, "fused", 3, 1, None)
        m = len(layer_passes(24, dtype, "fused"))
        assert launches == 1 + 3 * (m - 1), dtype


def _random_cases(n=20, seed=20251204):
    # qubit counts where the layers have 2 or 3 passes (fused N 14..24, depths
    # 1..6; half state N 24..25, depths 1..4)
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        dtype = ["c128", "c64"][i % 2]
        schedule = ["fused", "z2"][int(rng.integers(0, 2))]
        lo, hi = (24, 25) if schedule == "z2" else (14 if dtype == "c128" else 15, 24)
        N = int(rng.integers(lo, hi + 1))
        P = int(rng.integers(1, 5 if schedule == "z2" else 7))
        out.append((i, N, dtype, schedule, P))
    return out


@pytest.mark.parametrize("case", _random_cases(), ids=lambda c: f"{c[0]}-N{c[1]}-{c[2]}-{c[3]}-P{c[4]}")
def test_pipelined_random_vs_oracle(case, pipe_env):
    i, N, dtype, schedule, P = case
    print(f"case {i}: pipelined N={N} {dtype} {schedule} P={P}", flush=True)
    got, launches, edges, g_, b_ = _run(N, dtype, schedule, P, 500 + i, None)
    m = len(layer_passes(N, dtype, schedule))
    assert launches == (1 + P * (m - 1) if m >= (3 if schedule == "z2" else 2) else P * m)
    sv, qmap, W = O.run_weighted(N, 4, edges, g_, b_)
    ref = O.to_logical(sv, qmap)
    err = float(np.linalg.norm(got.astype(np.complex128) - ref) / np.linalg.norm(ref))
    print(f"  {launches} tile passes, relative L2 error vs oracle {err:.2e}", flush=True)
    assert err < TOL[dtype]
