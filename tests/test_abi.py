"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/forqaoa.h, validates arguments before touching the device,
and its host-side graph relabel matches the oracle's restatement of
mapping_after_swap_conti (src/swap.cu:27-50)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as O
from forqaoa import FQError, Graph, _lib
from forqaoa import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "forqaoa.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fq_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "fq_state_create" in syms and "fq_run" in syms and len(syms) >= 25


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [s for s in declared_symbols() if getattr(L, s, None) is None]
    assert not missing, missing


def test_every_declared_symbol_has_a_binding():
    missing = [s for s in declared_symbols() if s not in _lib.SIGNATURES]
    assert not missing, missing


def _create(N, D, B, C, dtype):
    h = ctypes.c_void_p()
    rc = _lib.lib().fq_state_create(N, D, B, C, dtype, ctypes.byref(h))
    return rc, h


@pytest.mark.parametrize(
    "args,code",
    [
        ((0, 0, 0, 1, 0), _lib.FQ_ERR_ARG),       # N < 1
        ((37, 0, 10, 10, 0), _lib.FQ_ERR_ARG),    # N too large
        ((20, 1, 10, 10, 0), _lib.FQ_ERR_UNSUPPORTED),  # D > 0 needs 2^D GPUs
        ((20, 0, 21, 10, 0), _lib.FQ_ERR_ARG),    # B > N - D (README.md:31)
        ((20, 0, 10, 14, 0), _lib.FQ_ERR_ARG),    # 2^14 complex128 tile > LDS
        ((20, 0, 10, 15, 1), _lib.FQ_ERR_ARG),    # 2^15 complex64 tile > LDS
        ((20, 0, 10, 0, 0), _lib.FQ_ERR_ARG),     # C < 1
        ((20, 0, 10, 10, 7), _lib.FQ_ERR_ARG),    # unknown dtype
    ],
)
def test_state_create_validates_before_allocating(args, code):
    rc, h = _create(*args)
    assert rc == code
    assert h.value is None
    assert _lib.lib().fq_last_error().decode()


def test_null_handles_are_rejected():
    L = _lib.lib()
    assert L.fq_cost_layer(None, None, 0.1, 0) == _lib.FQ_ERR_ARG
    assert L.fq_swap_conti(None, 0, 1, 1) == _lib.FQ_ERR_ARG
    assert L.fq_state_destroy(None) == _lib.FQ_OK
    assert L.fq_graph_destroy(None) == _lib.FQ_OK


def test_graph_add_edge_semantics():
    g = Graph(5)
    g.add_edge(3, 1, 0.5)
    g.add_edge(1, 3, 0.25)  # repeated edge overwrites (src/graph.cpp:63)
    g.add_edge(0, 4, 2.0)
    W = g.weights()
    assert W[1, 3] == 0.25 and W[3, 1] == 0 and W[0, 4] == 2.0
    with pytest.raises(_lib.FQError):
        g.add_edge(0, 5, 1.0)
    with pytest.raises(_lib.FQError):
        g.add_edge(2, 2, 1.0)
    with pytest.raises(_lib.FQError):
        g.add_edge(0, 1, float("nan"))


@pytest.mark.parametrize("N,out,inn,k", [(8, 0, 3, 3), (12, 0, 5, 5), (9, 1, 5, 4), (30, 0, 13, 13), (30, 0, 26, 4)])
def test_graph_relabel_matches_oracle(N, out, inn, k):
    edges = G.gnp(N, 0.6, 17 + N, "uniform")
    g = Graph(N, edges)
    W = O.dense_w(N, edges)
    g.relabel(out, inn, k)
    O.relabel(W, out, inn, k)
    np.testing.assert_array_equal(g.weights(), W)


def test_graph_relabel_rejects_overlap():
    g = Graph(8)
    with pytest.raises(_lib.FQError) as e:
        g.relabel(0, 2, 3)
    assert e.value.code == _lib.FQ_ERR_ARG


def test_driver_rejects_multi_device_like_reference_exit():
    exe = os.path.join(ROOT, "for-qaoa_amd", "host", "weighted")
    assert os.path.exists(exe), "driver not built"
    r = subprocess.run([exe, "1", "20", "1", "10", "10"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "D=1" in r.stderr
    r = subprocess.run([exe, "1", "20", "0", "10", "10", "--dtype", "c32"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2


def test_server_reports_requests_that_never_fit():
    # budget accounting happens before any device call, so this runs on the CPU
    from forqaoa.serve import Server, make_stream
    stream = make_stream(3, seed=5, n_range=(20, 20), p_range=(1, 1))
    srv = Server(dtype="c64", budget_bytes=1 << 20)
    try:
        with pytest.raises(FQError) as e:
            srv.run(stream, realtime=False)
        assert e.value.code == _lib.FQ_ERR_BUDGET
        assert "N=[20]" in str(e.value)
    finally:
        srv.close()
        _lib.check(_lib.lib().fq_set_memory_budget(0))


def test_stream_probe_validates_before_allocating():
    L = _lib.lib()
    d = ctypes.c_double()
    assert L.fq_stream_probe(1 << 10, 1, ctypes.byref(d), ctypes.byref(d), ctypes.byref(d)) == _lib.FQ_ERR_ARG
    assert L.fq_stream_probe((1 << 20) + 4, 1, ctypes.byref(d), ctypes.byref(d), ctypes.byref(d)) == _lib.FQ_ERR_ARG
    assert L.fq_stream_probe(1 << 20, 0, ctypes.byref(d), ctypes.byref(d), ctypes.byref(d)) == _lib.FQ_ERR_ARG
    assert L.fq_stream_probe(1 << 20, 1, None, None, None) == _lib.FQ_ERR_ARG
    _lib.check(L.fq_set_memory_budget(1 << 20))
    try:
        assert L.fq_stream_probe(2 << 20, 1, ctypes.byref(d), ctypes.byref(d), ctypes.byref(d)) == _lib.FQ_ERR_BUDGET
    finally:
        _lib.check(L.fq_set_memory_budget(0))


def _runs(pos):
    n, i = 0, 0
    while i < len(pos):
        j = i + 1
        while j < len(pos) and pos[j] == pos[j - 1] + 1:
            j += 1
        n, i = n + 1, j
    return n


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("schedule", ["fused", "z2"])
def test_layer_passes_cover_every_qubit_once(dtype, schedule):
    # host-only: the pass plan every fused / half-state layer runs
    from forqaoa import layer_passes
    qmax = 13 if dtype == "c128" else 14
    for N in range(3, 37):
        ps = layer_passes(N, dtype, schedule)
        Q = min(qmax, N - 1 if schedule == "z2" else N)
        assert ps[0][0] == list(range(Q)), (N, ps[0])
        seen = [q for _, g in ps for q in g]
        assert sorted(seen) == list(range(N)), (N, ps)        # one RX per qubit per layer
        a_log = max(1, min(3 if dtype == "c128" else 4, Q))
        for k, (tile, gates) in enumerate(ps):
            assert len(tile) == Q and len(set(tile)) == Q and set(gates) <= set(tile)
            if k > 0:
                c = next(i for i, q in enumerate(tile) if q != i) if tile != list(range(Q)) else Q
                assert tile[:c] == list(range(c))                # carriers: the lowest qubits first
            hidden = [p for p in range(N) if p not in tile]
            assert max(_runs(hidden), _runs(tile[:Q - a_log])) <= 8
        if schedule == "z2":
            # qubit N-1 is the top register bit of the last pass, hidden elsewhere
            assert ps[-1][0][-1] == N - 1
            assert all(N - 1 not in t for t, _ in ps[:-1])


def test_layer_passes_n30_c128():
    from forqaoa import layer_passes
    ps = layer_passes(30, "c128", "fused")
    assert len(ps) == 3
    assert ps[0][1] == list(range(5, 13))       # carried qubits 0..4 moved to the strided passes
    assert sorted(ps[1][1] + ps[2][1]) == [0, 1, 2, 3, 4] + list(range(13, 30))


@pytest.mark.parametrize("N,dtype,lanes", [
    (30, "c128", [[13, 14], [4, 20]]),   # 17 GiB state: lane bytes bits outside [19, 22]
    (29, "c128", [[4, 13], [4, 19]]),
    (32, "c128", [[14, 25], [26, 27]]),  # from 32 GiB on: outside [19, 28]
    (34, "c64", [[14, 15], [28, 29]]),
])
def test_layer_passes_wave_lanes(N, dtype, lanes):
    # tile positions 0..5 are the 64 lanes of a wave; past the carriers they
    # hold the lowest new qubits whose byte-address bit avoids the window
    from forqaoa import layer_passes
    ps = layer_passes(N, dtype, "fused")
    assert [t[4:6] for t, _ in ps[1:]] == lanes


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("schedule", ["fused", "z2"])
def test_boundary_pass_gets_no_carried_rx(dtype, schedule):
    # pipelined layers end alternately on pass 0 and on the last pass (fused)
    # or pass 1 (half state); with 3 or more passes per layer the carried low
    # qubits' RX goes to the other strided passes, which run it once per layer
    from forqaoa import layer_passes
    qmax = 13 if dtype == "c128" else 14
    for N in range(qmax + 2, 35):
        ps = layer_passes(N, dtype, schedule)
        if len(ps) < 3:
            continue
        q0 = len(ps[0][0])
        boundary = len(ps) - 1 if schedule == "fused" else 1

        def carriers(tile):
            c = 0
            while c < len(tile) and tile[c] == c:
                c += 1
            return c

        plain = [k for k in range(1, len(ps)) if k != boundary]
        # a low qubit's RX sits in the boundary pass only if no plain pass carries it
        for q in ps[boundary][1]:
            if q < q0:
                assert all(q >= carriers(ps[k][0]) for k in plain), (N, q, ps)
        assert any(q < q0 for k in plain for q in ps[k][1]), (N, ps)


@pytest.mark.parametrize("env", [{"FQ_TILE_Q": "10"}, {"FQ_TILE_Q": "11"}, {"FQ_TILE_Q": "12"},
                                 {"FQ_PASS_BLOCK": "0"}, {"FQ_PASS_BLOCK": "1"}, {"FQ_PASS_BLOCK": "3"},
                                 {"FQ_PASS_BLOCK": "4"}, {"FQ_CARRIER_GATES": "0"}])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("schedule", ["fused", "z2"])
def test_layer_passes_under_tuning_knobs(env, dtype, schedule, monkeypatch):
    # host-only: the knobs (read on every call) still give one RX per qubit per
    # layer, tiles of distinct positions, and the mirror qubit last
    from forqaoa import layer_passes
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    qmax = min(13 if dtype == "c128" else 14, int(env.get("FQ_TILE_Q", "99")))
    for N in range(3, 37):
        ps = layer_passes(N, dtype, schedule)
        Q = min(qmax, N - 1 if schedule == "z2" else N)
        assert ps[0][0] == list(range(Q)), (N, ps[0])
        assert sorted(q for _, g in ps for q in g) == list(range(N)), (N, ps)
        for tile, gates in ps:
            assert len(tile) == Q and len(set(tile)) == Q and set(gates) <= set(tile)
        if env.get("FQ_CARRIER_GATES") == "0":
            assert set(range(Q)) <= set(ps[0][1]), (N, ps)
        if schedule == "z2":
            assert ps[-1][0][-1] == N - 1
