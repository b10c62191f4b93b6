"""Adjoint gradient (fq_backward) against finite differences.

* small N: the complex128 adjoint against central differences (h = 1e-4) of
  the CPU oracle's <C> -- an independent forward path; tolerance 1e-7 of
  max(1, |g|) (the central difference itself is accurate to ~1e-9 here);
  the complex64 adjoint within 1e-4.
* the adjoint workload (N=26, P=8, seeded 4-regular graph, weights in (0,1],
  16 random angle sets): complex64 adjoint against central differences
  (h = 1e-4) of the complex128 GPU forward, and against complex128 forward
  differences (h = 1e-4, truncation O(h)); the complex128 adjoint against the
  same central differences.
* the torch.autograd.Function wrapper returns the same gradients.
"""
import numpy as np
import pytest

import pyoracle as O
from forqaoa import FQError, Graph, State
from forqaoa import graphs as G

pytestmark = pytest.mark.gpu


def oracle_expect(N, edges, g, b):
    sv, qmap, W = O.run_weighted(N, 4, edges, g, b)
    return O.expect_cut(sv, W)


def central_fd(f, g, b, h):
    P = len(g)
    gg = np.zeros(P)
    gb = np.zeros(P)
    for k in range(P):
        gp, gm = g.copy(), g.copy()
        gp[k] += h
        gm[k] -= h
        gg[k] = (f(gp, b) - f(gm, b)) / (2 * h)
        bp, bm = b.copy(), b.copy()
        bp[k] += h
        bm[k] -= h
        gb[k] = (f(g, bp) - f(g, bm)) / (2 * h)
    return gg, gb


@pytest.mark.parametrize("N,P,seed", [(4, 2, 5), (5, 3, 6), (6, 2, 1), (10, 3, 2), (13, 2, 3), (16, 3, 4)])
@pytest.mark.parametrize("dtype,tol", [("c128", 1e-7), ("c64", 1e-4)])
def test_adjoint_vs_oracle_central_difference(N, P, seed, dtype, tol):
    edges = G.gnp(N, 0.5, 900 + seed, "uniform")
    g, b = G.angles(P, 900 + seed)
    st = State(N, C=min(N, 10), dtype=dtype)
    gr = Graph(N, edges)
    st.run(gr, g, b, "fused")
    e, gg, gb = st.backward(gr, g, b)
    e_ref = oracle_expect(N, edges, g, b)
    assert e == pytest.approx(e_ref, rel=1e-12 if dtype == "c128" else 1e-5)
    fg, fb = central_fd(lambda x, y: oracle_expect(N, edges, x, y), g, b, 1e-4)
    scale = max(1.0, float(np.max(np.abs(np.concatenate([fg, fb])))))
    err = float(np.max(np.abs(np.concatenate([gg - fg, gb - fb])))) / scale
    print(f"adjoint {dtype} N={N} P={P}: max |g_adj - g_fd| / max(1,|g|) = {err:.2e}", flush=True)
    assert err < tol


def test_adjoint_workload_n26_p8_reg4():
    w = G.workload("n26_reg4")
    N, P = w["N"], w["P"]
    gr = Graph(N, w["edges"])
    st128 = State(N, C=13, dtype="c128")
    st64 = State(N, C=14, dtype="c64")

    def f128(x, y):
        st128.run(gr, x, y, "fused")
        return st128.expect_cut(gr)[0]

    rng = np.random.default_rng(2026)
    worst_c = worst_f = worst_128 = 0.0
    for s in range(16):
        g = rng.uniform(0, np.pi, P)
        b = rng.uniform(0, np.pi, P)
        st64.run(gr, g, b, "fused")
        e64, gg64, gb64 = st64.backward(gr, g, b)
        st128.run(gr, g, b, "fused")
        e128, gg128, gb128 = st128.backward(gr, g, b)
        fg, fb = central_fd(f128, g, b, 1e-4)
        e0 = f128(g, b)
        fwd_g = np.array([(f128(g + 1e-4 * np.eye(P)[k], b) - e0) / 1e-4 for k in range(P)])
        fwd_b = np.array([(f128(g, b + 1e-4 * np.eye(P)[k]) - e0) / 1e-4 for k in range(P)])
        scale = max(1.0, float(np.max(np.abs(np.concatenate([fg, fb])))))
        worst_c = max(worst_c, float(np.max(np.abs(np.concatenate([gg64 - fg, gb64 - fb])))) / scale)
        worst_128 = max(worst_128, float(np.max(np.abs(np.concatenate([gg128 - fg, gb128 - fb])))) / scale)
        worst_f = max(worst_f, float(np.max(np.abs(np.concatenate([gg64 - fwd_g, gb64 - fwd_b])))) / scale)
        assert e64 == pytest.approx(e128, rel=1e-5)
    print(f"N=26 P=8 reg4, 16 angle sets: c64 adjoint vs c128 central FD {worst_c:.2e}; "
          f"c128 adjoint vs central FD {worst_128:.2e}; c64 adjoint vs c128 forward FD {worst_f:.2e}",
          flush=True)
    assert worst_128 < 1e-7
    assert worst_c < 1e-4
    assert worst_f < 1e-2  # forward differences carry an O(h) truncation error


def test_backward_requires_identity_layout():
    N = 10
    edges = G.gnp(N, 0.5, 3, "uniform")
    g, b = G.angles(2, 3)
    st = State(N, C=4)
    gr = Graph(N, edges)
    st.run(gr, g, b, "reference")
    assert list(st.qubit_map()) != list(range(N))
    with pytest.raises(FQError):
        st.backward(gr, g, b)


def test_torch_autograd_function():
    import torch

    from forqaoa.autograd import qaoa_expectation

    N, P = 12, 3
    edges = G.gnp(N, 0.5, 77, "uniform")
    g, b = G.angles(P, 77)
    gr = Graph(N, edges)
    gt = torch.tensor(g, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
    E = qaoa_expectation(gt, bt, gr, dtype="c128")
    E.backward()
    st = State(N, C=10)
    st.run(gr, g, b, "fused")
    e, gg, gb = st.backward(gr, g, b)
    assert float(E.detach()) == pytest.approx(e, rel=1e-13)
    np.testing.assert_allclose(gt.grad.numpy(), gg, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(bt.grad.numpy(), gb, rtol=1e-12, atol=1e-13)
    # one descent step on -<C> increases the expected cut
    with torch.no_grad():
        g2 = gt + 0.01 * gt.grad
        b2 = bt + 0.01 * bt.grad
    E2 = qaoa_expectation(g2, b2, gr, dtype="c128")
    assert float(E2) > float(E)


def _random_adjoint_cases(n=16, seed=20251203):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        N = int(rng.integers(3, 19))
        P = int(rng.integers(1, 5))
        dtype = ["c128", "c64"][i % 2]
        p = float(rng.uniform(0.3, 0.9))
        kind = ["uniform", "int16", "unit"][int(rng.integers(0, 3))]
        out.append((i, N, P, dtype, p, kind))
    return out


@pytest.mark.parametrize("case", _random_adjoint_cases(), ids=lambda c: f"{c[0]}-N{c[1]}-P{c[2]}-{c[3]}")
def test_adjoint_random_configurations(case):
    # seeded graphs and depths across the tile-width boundaries: the half-state
    # walk back (mirror pass, doubled sums) against central differences of the
    # oracle's <C>
    i, N, P, dtype, p, kind = case
    print(f"case {i}: N={N} P={P} {dtype} G(N,{p:.2f}) {kind}", flush=True)
    edges = G.gnp(N, p, 500 + i, kind)
    g, b = G.angles(P, 600 + i)
    st = State(N, C=min(N, 10), dtype=dtype)
    gr = Graph(N, edges)
    try:
        st.run(gr, g, b, "z2" if N >= 3 else "fused")
        e, gg, gb = st.backward(gr, g, b)
        scale_e = max(1.0, abs(oracle_expect(N, edges, g, b)))
        assert abs(e - oracle_expect(N, edges, g, b)) / scale_e < (1e-12 if dtype == "c128" else 1e-5)
        # the phase angles scale with the weights (integer weights up to 16):
        # the difference step shrinks with them, so the central difference's
        # O(h^2) truncation stays below the tolerance
        wmax = max([1.0] + [abs(w) for _, _, w in edges])
        fg, fb = central_fd(lambda x, y: oracle_expect(N, edges, x, y), g, b, 1e-4 / wmax)
        scale = max(1.0, float(np.max(np.abs(np.concatenate([fg, fb])))))
        err = float(np.max(np.abs(np.concatenate([gg - fg, gb - fb])))) / scale
        print(f"  max |g_adj - g_fd| / max(1,|g|) = {err:.2e}", flush=True)
        assert err < (1e-6 if dtype == "c128" else 1e-4)
    finally:
        st.destroy()
        gr.destroy()


@pytest.mark.parametrize("N,P", [(21, 2), (22, 3), (24, 4), (23, 5)])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_pipelined_walk_back_matches_layer_order(N, P, dtype, monkeypatch):
    # From 3 pair passes per layer on, the walk back undoes the launches of a
    # pipelined forward: layer-boundary pair passes undo RX(beta_{p+1}), the
    # phase gamma_{p+1} and RX(beta_p) in one launch.  The gradients equal the
    # layer-order walk back's (FQ_LAYER_PIPE=0) and central differences.
    edges = G.gnp(N, 0.4, 700 + N, "uniform")
    g, b = G.angles(P, 710 + N)
    gr = Graph(N, edges)
    st = State(N, C=min(N, 13 if dtype == "c128" else 14), dtype=dtype)
    try:
        st.run(gr, g, b, "z2")
        e1, gg1, gb1 = st.backward(gr, g, b)
        monkeypatch.setenv("FQ_LAYER_PIPE", "0")
        st.run(gr, g, b, "z2")
        e0, gg0, gb0 = st.backward(gr, g, b)
        monkeypatch.delenv("FQ_LAYER_PIPE")
        ref = State(N, C=13, dtype="c128")

        def f(x, y):
            ref.run(gr, x, y, "fused")
            return ref.expect_cut(gr)[0]

        try:
            fg, fb = central_fd(f, g, b, 1e-4)
        finally:
            ref.destroy()
    finally:
        st.destroy()
        gr.destroy()
    scale = max(1.0, float(np.max(np.abs(np.concatenate([gg0, gb0])))))
    d_order = float(np.max(np.abs(np.concatenate([gg1 - gg0, gb1 - gb0])))) / scale
    d_fd = float(np.max(np.abs(np.concatenate([gg1 - fg, gb1 - fb])))) / scale
    print(f"walk back N={N} P={P} {dtype}: pipelined vs layer order {d_order:.2e}, "
          f"vs central FD {d_fd:.2e}", flush=True)
    assert e1 == pytest.approx(e0, rel=1e-12 if dtype == "c128" else 1e-5)
    assert d_order < (1e-12 if dtype == "c128" else 1e-5)
    assert d_fd < (1e-7 if dtype == "c128" else 1e-4)
