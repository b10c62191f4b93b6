"""Packed slabs and the continuous-batching server.

Each request's <C> in the mix must equal the same request run alone through
fq_run (fused schedule, its own state) and the oracle's <C> of the same
request (the reference's layer loop, src/Weighted.cpp:97-137, where the
request is small enough to run it here): within 1e-5 relative for complex64
(the stated tolerance of the complex64 path) and 1e-12 for complex128; a
slot's full state equals the solo state within the same tolerance.  A slice
of the serving workload itself (N 16..30, P 1..10, 200 GB budget) runs
through the server with its N=29 and N=30 requests checked against solo runs
and complex128.
"""
import ctypes

import numpy as np
import pytest

import pyoracle as O
from forqaoa import FQError, Graph, State, _lib
from forqaoa import graphs as G
from forqaoa.serve import Server, Slab, in_use, latency_stats, make_stream, one_at_a_time

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


def solo(N, edges, g, b, dtype):
    st = State(N, C=min(N, 13 if dtype == "c128" else 14), dtype=dtype)
    gr = Graph(N, edges)
    st.run(gr, g, b, "fused")
    e, _ = st.expect_cut(gr)
    amps = st.copy_out()
    st.destroy()
    return e, amps


@pytest.mark.parametrize("schedule", ["fused", "z2"])
@pytest.mark.parametrize("dtype", ["c64", "c128"])
@pytest.mark.parametrize("N", [15, 21, 25])  # 25: each slot runs its own tile-pass launches
def test_slab_slots_match_solo_runs(dtype, schedule, N):
    sb = Slab(N, 8, dtype, schedule)
    jobs = []
    for k in range(5):
        edges = G.gnp(N, 0.3 + 0.1 * k, 500 + k, "uniform")
        P = 1 + k
        g, b = G.angles(P, 600 + k)
        gr = Graph(N, edges)
        jobs.append((k, gr, edges, g, b))
        sb.assign(k, gr, type("R", (), {"gammas": g, "betas": b})())
    # step until every slot finished; slot k has P = k + 1 layers
    results = {}
    amps = {}
    for step in range(5):
        sb.step()
        sb.poll()
        # copy each slot out right when it finishes (before it is collected)
        for k, gr, edges, g, b in jobs:
            if len(g) == step + 1:
                amps[k] = sb.copy_out(k, dtype)
        for r in sb.collect():
            pass
    for k, gr, edges, g, b in jobs:
        e_solo, a_solo = solo(N, edges, g, b, dtype)
        err = np.linalg.norm(amps[k].astype(np.complex128) - a_solo) / np.linalg.norm(a_solo)
        assert err < TOL[dtype], (k, err)
    sb.destroy()


_ORACLE = {}


def oracle_expect(N, edges, g, b):
    key = (N, tuple(edges), tuple(g), tuple(b))
    if key not in _ORACLE:
        sv, _, W = O.run_weighted(N, min(N, 10), edges, g, b)
        _ORACLE[key] = O.expect_cut(sv, W)
    return _ORACLE[key]


@pytest.mark.parametrize("schedule", ["z2", "fused"])
@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_server_results_equal_solo(dtype, schedule):
    stream = make_stream(60, seed=11, n_range=(3, 18), p_range=(1, 4), rate=None)
    # budgets count from what this process already holds
    srv = Server(dtype=dtype, budget_bytes=in_use() + 8e9, schedule=schedule)
    done = srv.run(stream, realtime=False)
    assert len(done) == 60
    worst = worst_o = 0.0
    for r in done:
        N, edges = stream.graphs[r.graph_key]
        e_solo, _ = solo(N, edges, r.gammas, r.betas, dtype)
        e_or = oracle_expect(N, edges, list(r.gammas), list(r.betas))
        worst = max(worst, abs(r.expect - e_solo) / abs(e_solo))
        worst_o = max(worst_o, abs(r.expect - e_or) / abs(e_or))
        assert r.norm2 == pytest.approx(1.0, abs=1e-5 if dtype == "c64" else 1e-11)
    print(f"server {dtype} {schedule}: max relative |<C>_mix - <C>_alone| = {worst:.2e}, "
          f"|<C>_mix - <C>_oracle| = {worst_o:.2e}", flush=True)
    assert worst < TOL[dtype]
    assert worst_o < TOL[dtype]
    # requests joined slabs that already had requests in flight
    assert any(r.joined_running > 0 for r in done)
    srv.close()


def test_server_continuous_admission_realtime():
    # arrivals while earlier requests run: later requests start before earlier ones finish
    stream = make_stream(80, seed=12, n_range=(14, 20), p_range=(2, 8), rate=4000.0)
    srv = Server(dtype="c64", budget_bytes=in_use() + 8e9)
    done = srv.run(stream, realtime=True)
    assert len(done) == 80
    overlaps = sum(1 for a in done for b in done if a.rid < b.rid and b.start < a.done and a.N == b.N)
    assert overlaps > 0
    st = latency_stats(done)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()


def test_budget_limits_admission_and_server_still_finishes():
    # a budget that holds one slab at a time: slabs are released when idle
    stream = make_stream(24, seed=13, n_range=(16, 19), p_range=(1, 3), rate=None)
    base = in_use()
    srv = Server(dtype="c64", budget_bytes=base + 150e6, slots_for=lambda N: 2)
    done = srv.run(stream, realtime=False)
    assert len(done) == 24
    assert in_use() <= base + 150e6
    srv.close()
    _lib.check(_lib.lib().fq_set_memory_budget(0))


def test_budget_error_code():
    _lib.check(_lib.lib().fq_set_memory_budget(in_use() + (1 << 20)))
    try:
        with pytest.raises(FQError) as e:
            State(20, C=10, dtype="c64")
        assert e.value.code == _lib.FQ_ERR_BUDGET
    finally:
        _lib.check(_lib.lib().fq_set_memory_budget(0))


def test_one_at_a_time_baseline_agrees():
    stream = make_stream(6, seed=14, n_range=(12, 14), p_range=(1, 3))
    srv = Server(dtype="c128", budget_bytes=in_use() + 4e9)
    done = {r.rid: r.expect for r in srv.run(stream, realtime=False)}
    srv.close()
    one_at_a_time(stream, stream.requests, dtype="c128", schedule="reference")
    for r in stream.requests:
        assert r.expect == pytest.approx(done[r.rid], rel=1e-12)


def _random_slab_cases(n=12, seed=20251204):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        N = int(rng.integers(3, 25))
        K = int([2, 4, 8][int(rng.integers(0, 3))])
        schedule = ["fused", "z2"][int(rng.integers(0, 2))]
        dtype = ["c128", "c64"][i % 2]
        out.append((i, N, K, schedule, dtype))
    return out


@pytest.mark.parametrize("case", _random_slab_cases(), ids=lambda c: f"{c[0]}-N{c[1]}-K{c[2]}-{c[3]}-{c[4]}")
def test_slab_random_sizes_match_solo(case):
    # seeded qubit counts, slot counts and depths: every request's <C> from a
    # packed slab equals the same request run alone
    i, N, K, schedule, dtype = case
    print(f"case {i}: N={N} K={K} {schedule} {dtype}", flush=True)
    rng = np.random.default_rng(700 + i)
    sb = Slab(N, K, dtype, schedule)
    graphs = []
    try:
        reqs = []
        for k in range(min(K, 3)):
            edges = G.gnp(N, float(rng.uniform(0.3, 0.9)), 800 + 10 * i + k, "uniform")
            P = int(rng.integers(1, 5))
            g, b = G.angles(P, 900 + 10 * i + k)
            gr = Graph(N, edges)
            graphs.append(gr)
            r = type("R", (), {"gammas": g, "betas": b, "edges": edges, "expect": None})()
            reqs.append(r)
            sb.assign(k, gr, r)
        for _ in range(8):
            if all(r.expect is not None for r in reqs):
                break
            sb.step()
            sb.poll()
            sb.collect()
        for r in reqs:
            e_solo, _ = solo(N, r.edges, r.gammas, r.betas, dtype)
            assert r.expect == pytest.approx(e_solo, rel=TOL[dtype], abs=TOL[dtype])
    finally:
        sb.destroy()
        for gr in graphs:
            gr.destroy()


def test_serving_workload_slice_under_budget():
    # every 25th request of the serving workload's own stream (N 16..30,
    # P 1..10, G(N, p) p in [0.2, 0.8], 80% resident-graph re-evaluations):
    # 400 requests through half-state complex64 slabs under a 200 GB budget
    full = make_stream(10000, seed=20251207)
    reqs = full.requests[::25]
    for r in reqs:
        r.arrival = 0.0
    stream = type(full)(graphs=full.graphs, requests=reqs)
    base = in_use()
    budget = base + 200e9
    srv = Server(dtype="c64", budget_bytes=budget)
    try:
        done = srv.run(stream, realtime=False)
        peak = srv.peak_in_use
    finally:
        srv.close()
        _lib.check(_lib.lib().fq_set_memory_budget(0))
    assert len(done) == len(reqs)
    worst_norm = max(abs(r.norm2 - 1.0) for r in done)
    print(f"slice: {len(done)} requests, N {min(r.N for r in done)}..{max(r.N for r in done)}, "
          f"peak HBM {(peak - base) / 1e9:.1f} GB of 200, max |norm2 - 1| {worst_norm:.2e}", flush=True)
    assert peak <= budget
    assert worst_norm < 1e-5
    # the mix against solo runs: every N=29 and N=30 request (up to 4 each),
    # and one request of each other size; the N=30 ones also against complex128
    by_n = {}
    for r in done:
        by_n.setdefault(r.N, []).append(r)
    check = []
    for N, rs in sorted(by_n.items()):
        check += rs[:4] if N >= 29 else rs[:1]
    assert any(r.N == 30 for r in check) and any(r.N == 29 for r in check)
    worst = worst128 = 0.0
    for r in check:
        N, edges = full.graphs[r.graph_key]
        e_solo = solo_expect(N, edges, r.gammas, r.betas, "c64")
        worst = max(worst, abs(r.expect - e_solo) / abs(e_solo))
        if N == 30:
            e128 = solo_expect(N, edges, r.gammas, r.betas, "c128")
            worst128 = max(worst128, abs(r.expect - e128) / abs(e128))
    print(f"slice: {len(check)} requests vs solo runs: max rel diff {worst:.2e}; "
          f"N=30 complex64 mix vs complex128 solo: {worst128:.2e}", flush=True)
    assert worst < TOL["c64"]
    assert worst128 < TOL["c64"]


def solo_expect(N, edges, g, b, dtype):
    st = State(N, C=min(N, 13 if dtype == "c128" else 14), dtype=dtype)
    gr = Graph(N, edges)
    try:
        st.run(gr, g, b, "fused")
        return st.expect_cut(gr)[0]
    finally:
        st.destroy()
        gr.destroy()


@pytest.mark.parametrize("layers", [2, 3, 10])
@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_large_slots_several_layers_per_step(dtype, layers):
    # slots of 2^10 or more first-pass tiles run fq_run's passes per slot; a
    # step may run several layers of each (fewer for a slot with fewer left)
    N = 25
    sb = Slab(N, 4, dtype, "z2")
    sb.set_layers_per_step(layers)
    jobs = {}
    try:
        for k in range(4):
            edges = G.gnp(N, 0.25 + 0.15 * k, 900 + k, "uniform")
            P = 2 + 2 * k
            g, b = G.angles(P, 910 + k)
            gr = Graph(N, edges)
            jobs[k] = (gr, edges, g, b)
            sb.assign(k, gr, type("R", (), {"gammas": g, "betas": b})())
        got = {}
        steps = 0
        while len(got) < 4:
            sb.step()
            steps += 1
            for r in sb.collect():
                got[len(got)] = r
        assert steps == -(-8 // layers)   # the longest request has P = 8
        by_p = {len(r.gammas): r for r in got.values()}
        worst = 0.0
        for k, (gr, edges, g, b) in jobs.items():
            e_solo = solo_expect(N, edges, g, b, dtype)
            r = by_p[len(g)]
            worst = max(worst, abs(r.expect - e_solo) / abs(e_solo))
            assert r.norm2 == pytest.approx(1.0, abs=1e-5 if dtype == "c64" else 1e-11)
        print(f"N=25 {dtype} {layers} layers per step: max relative |<C>_slab - <C>_alone| = {worst:.2e}",
              flush=True)
        assert worst < TOL[dtype]
    finally:
        sb.destroy()
        for gr, *_ in jobs.values():
            gr.destroy()
