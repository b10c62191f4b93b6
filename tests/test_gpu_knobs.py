"""Every tuning knob of the fused and half-state schedules against the oracle.

The library reads these environment variables on every call; each selects a
different tile layout or launch order, so each is a separate product path:

* FQ_TILE_Q      tile width of the fused and strided passes (below the LDS
                 limit of 13 qubits complex128 / 14 complex64);
* FQ_TILE_ORDER  order in which the persistent work-groups walk the tiles
                 (0 sweeps, 1 XCD-interleaved, 3 one work-group per tile;
                 2 is the default);
* FQ_PASS_BLOCK  block size in which the high qubits are dealt to the strided
                 passes (0: one contiguous block per pass; 2 is the default);
* FQ_CARRIER_GATES=0  the carried low qubits' RX stays in the cost pass;
* FQ_WIDE_REGS=0  the contiguous 14-qubit complex64 tiles run 1024 threads
                 x 16 amplitudes instead of 512 x 32.

States in logical order within 1e-12 (complex128) / 1e-5 (complex64) of the
oracle's final state, at sizes where the strided passes exist (N > tile
width), for pipelined layers and layer order.
"""
import os

import numpy as np
import pytest

import pyoracle as O
from forqaoa import Graph, State
from forqaoa import graphs as G

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}
KNOBS = ["FQ_TILE_Q", "FQ_TILE_ORDER", "FQ_PASS_BLOCK", "FQ_CARRIER_GATES", "FQ_LAYER_PIPE", "FQ_WIDE_REGS"]

CASES = [
    ({"FQ_TILE_Q": "10"}, 22, "c128"),
    ({"FQ_TILE_Q": "12"}, 25, "c64"),
    ({"FQ_TILE_Q": "11", "FQ_LAYER_PIPE": "0"}, 23, "c128"),
    ({"FQ_TILE_ORDER": "0"}, 23, "c128"),
    ({"FQ_TILE_ORDER": "1"}, 24, "c64"),
    ({"FQ_TILE_ORDER": "3"}, 22, "c128"),
    ({"FQ_PASS_BLOCK": "0"}, 25, "c128"),
    ({"FQ_PASS_BLOCK": "1"}, 24, "c64"),
    ({"FQ_PASS_BLOCK": "3"}, 25, "c64"),
    ({"FQ_PASS_BLOCK": "4", "FQ_LAYER_PIPE": "0"}, 24, "c128"),
    ({"FQ_CARRIER_GATES": "0"}, 23, "c128"),
    ({"FQ_CARRIER_GATES": "0"}, 24, "c64"),
    ({"FQ_CARRIER_GATES": "0", "FQ_LAYER_PIPE": "0"}, 22, "c64"),
    # 512 x 32 is the default where it saves an LDS phase of the contiguous
    # pass: at N=23 (and 32, 33) its RX set (qubits 5..13) takes one phase of
    # five bits instead of two of four; at N=24 (and 34, RX on 4..13) only the
    # write-only layer-0 pass keeps it
    ({"FQ_WIDE_REGS": "1"}, 24, "c64"),
    ({"FQ_WIDE_REGS": "1", "FQ_LAYER_PIPE": "0"}, 23, "c64"),
    ({"FQ_WIDE_REGS": "0"}, 24, "c64"),
    ({"FQ_WIDE_REGS": "0", "FQ_LAYER_PIPE": "0"}, 23, "c64"),
]


@pytest.fixture
def clean_env():
    old = {k: os.environ.get(k) for k in KNOBS}
    yield
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


@pytest.mark.parametrize("schedule", ["fused", "z2"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(f"{k}={v}" for k, v in c[0].items()) + f"-N{c[1]}-{c[2]}")
def test_knob_vs_oracle(clean_env, case, schedule):
    env, N, dtype = case
    i = CASES.index(case)
    print(f"knobs {env} N={N} {dtype} {schedule}", flush=True)
    edges = G.gnp(N, 0.5, 7000 + i, ["uniform", "int16", "unit"][i % 3])
    gm, bt = G.angles(3, 7100 + i)
    sv, qmap, W = O.run_weighted(N, min(N, 10), edges, gm, bt)
    want = O.to_logical(sv, qmap)
    for k in KNOBS:
        os.environ.pop(k, None)
    os.environ.update(env)
    st = State(N, C=min(N, 10), dtype=dtype)
    g = Graph(N, edges)
    try:
        st.run(g, gm, bt, schedule)
        got = st.copy_out(logical=True).astype(np.complex128)
        err = float(np.linalg.norm(got - want) / np.linalg.norm(want))
        ex, _ = st.expect_cut(g)
    finally:
        st.destroy()
        g.destroy()
    rel = abs(ex - O.expect_cut(sv, W)) / abs(O.expect_cut(sv, W))
    print(f"  state rel err {err:.2e}, <C> rel err {rel:.2e}", flush=True)
    assert err < TOL[dtype]
    assert rel < TOL[dtype]
