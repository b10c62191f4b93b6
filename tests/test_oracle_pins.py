"""Pins for every function of the CPU oracle (oracle/fq_oracle.c).

The reference ships no tests, fixtures or golden vectors for this path
(its only check, scripts/valiate_ans_scripts/validate_state.py, needs
qiskit-aer-gpu and a state dump the code never writes).  Each oracle function
is therefore pinned by mathematics that is independent of its own formulas:

function             pin
-------------------  ---------------------------------------------------------
oq_init_amp          2^(-N/2) closed form
oq_flatten_upper     brute-force enumeration order
oq_cost_layer        full-matrix RZZ(w*gamma) gates (dense simulator); exact
                     integer s(b); |psi|^2 = 2^-N; complement symmetry
oq_multi_rx          full-matrix RX(beta) = expm(-i beta/2 X) gates; chunking
oq_swap_conti        numpy axis permutation (bitwise); involution
oq_relabel           brute force: Cut on the relabelled W at permuted indices
oq_run_weighted      dense simulator on non-symmetric graphs (pins layout,
                     qubit map and schedule); Hamming-weight closed form (K_N);
                     analytic single-layer <C>
oq_expect_cut        Z-string cut diagonal; closed-form <C> for K_N
oq_norm2, oq_topk    numpy
oq_to_logical        numpy axis transposition
oq_cut_value         Z-string cut diagonal
"""
import math

import numpy as np
import pytest

import golden_util as GU
import pyoracle as O
import refmath as R
from forqaoa import graphs as G


def test_init_amp_closed_form():
    for N in range(1, 37):
        assert O.init_amp(N) == pytest.approx(2.0 ** (-N / 2), rel=1e-14)


def test_flatten_order():
    N = 6
    W = np.array([[100 * i + j for j in range(N)] for i in range(N)], dtype=np.float64)
    flat = np.zeros(N * (N - 1) // 2)
    import ctypes

    O.lib().oq_flatten_upper.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]
    O.lib().oq_flatten_upper(N, O._dp(np.ascontiguousarray(W)), O._dp(flat))
    expect = [100 * i + j for i in range(N) for j in range(i + 1, N)]
    assert list(flat) == expect


@pytest.mark.parametrize("N,seed", [(4, 1), (6, 2), (8, 3)])
def test_cost_layer_vs_dense_rzz(N, seed):
    edges = G.gnp(N, 0.6, seed, "uniform")
    gamma = 1.234
    sv = np.zeros(1 << N, dtype=np.complex128)
    O.cost_layer(sv, O.dense_w(N, edges), gamma, True)
    # dense: H on all then RZZ(w*gamma) per edge, no mixer
    psi = R.dense_qaoa(N, edges, [gamma], [0.0])
    np.testing.assert_allclose(sv, psi, rtol=0, atol=1e-15)
    # second (read-modify-write) layer on top
    O.cost_layer(sv, O.dense_w(N, edges), -0.77, False)
    psi2 = R.dense_qaoa(N, edges, [gamma + -0.77], [0.0])
    np.testing.assert_allclose(sv, psi2, rtol=0, atol=1e-15)


def test_cost_layer_integer_weights_exact_phase():
    N = 10
    edges = G.gnp(N, 0.5, 11, "int16")
    W = O.dense_w(N, edges)
    gamma = 0.5
    sv = np.zeros(1 << N, dtype=np.complex128)
    O.cost_layer(sv, W, gamma, True)
    amp = 2.0 ** (-N / 2)
    np.testing.assert_allclose(np.abs(sv) ** 2, np.full(1 << N, amp * amp), rtol=1e-14)
    # s(b) = 2 Cut(b) - W exactly; phase = s*gamma/2
    cut = R.dense_cut_diag(N, edges)
    s = 2 * cut - W.sum()
    assert np.all(s == np.round(s))
    np.testing.assert_allclose(sv, amp * np.exp(1j * s * gamma / 2), atol=1e-15)
    # complement symmetry psi(b) = psi(~b)
    comp = (np.arange(1 << N) ^ ((1 << N) - 1))
    np.testing.assert_array_equal(sv, sv[comp])


def test_multi_rx_vs_dense_rx():
    N, C = 7, 4
    rng = np.random.default_rng(5)
    sv = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    ref = sv.copy()
    beta = 0.9
    targets = [0, 2, 3]
    O.multi_rx(sv, N, C, targets, beta)
    import scipy.linalg

    RX = scipy.linalg.expm(-1j * beta / 2 * R._X)
    for q in targets:
        ref = R._embed(N, {q: RX}) @ ref
    np.testing.assert_allclose(sv, ref, atol=1e-14)


def test_multi_rx_rejects_targets_outside_chunk():
    sv = np.zeros(1 << 6, dtype=np.complex128)
    with pytest.raises(ValueError):
        O.multi_rx(sv, 6, 3, [3], 0.1)


@pytest.mark.parametrize("N,out,inn,k", [(6, 0, 3, 3), (7, 0, 4, 2), (8, 1, 4, 3), (9, 0, 3, 1), (10, 2, 5, 3), (10, 1, 6, 4)])
def test_swap_vs_axis_permutation(N, out, inn, k):
    rng = np.random.default_rng(N * 10 + k)
    sv = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    ref = R.axis_swap_blocks(sv, N, out, inn, k)
    a = sv.copy()
    O.swap_conti(a, N, out, inn, k)
    np.testing.assert_array_equal(a, ref)
    O.swap_conti(a, N, out, inn, k)
    np.testing.assert_array_equal(a, sv)


def test_swap_rejects_overlap():
    sv = np.zeros(1 << 6, dtype=np.complex128)
    with pytest.raises(ValueError):
        O.swap_conti(sv, 6, 0, 2, 3)


@pytest.mark.parametrize("N,out,inn,k", [(8, 0, 3, 3), (9, 1, 5, 4), (7, 0, 5, 2)])
def test_relabel_brute_force(N, out, inn, k):
    edges = G.gnp(N, 0.7, 3 + N, "uniform")
    W = O.dense_w(N, edges)
    W2 = W.copy()
    O.relabel(W2, out, inn, k)

    def perm(b):
        for t in range(k):
            a, c = out + t, inn + t
            ba, bc = (b >> a) & 1, (b >> c) & 1
            if ba != bc:
                b ^= (1 << a) | (1 << c)
        return b

    for b in range(1 << N):
        assert O.cut_value(W2, perm(b)) == pytest.approx(O.cut_value(W, b), abs=1e-12)
    O.relabel(W2, out, inn, k)
    np.testing.assert_array_equal(W2, W)


def test_run_vs_dense_fixtures():
    inst = GU.instances()
    dense = GU.dense_small()
    for name, psi in dense.items():
        d = inst[name]
        sv, qmap, W = O.run_weighted(d["N"], d["C"], d["edges"], d["gammas"], d["betas"])
        lg = O.to_logical(sv, qmap)
        np.testing.assert_allclose(lg, psi, rtol=0, atol=1e-14, err_msg=name)
        # the final relabelled W is consistent with the physical layout
        e_phys = O.expect_cut(sv, W)
        e_dense = float(np.sum(np.abs(psi) ** 2 * R.dense_cut_diag(d["N"], d["edges"])))
        assert e_phys == pytest.approx(e_dense, rel=1e-13), name


def test_run_layout_is_permuted_as_traced():
    # final physical->logical maps traced from src/Weighted.cpp:97-137
    for (N, C, P), expect in [((20, 12, 5), list(range(12, 20)) + list(range(8, 12)) + list(range(0, 8))),
                              ((12, 5, 1), list(range(10, 12)) + list(range(2, 5)) + list(range(0, 2))
                               + list(range(5, 10)))]:
        edges = G.gnp(N, 0.3, 1, "uniform")
        g, b = G.angles(P, 2)
        _, qmap, _ = O.run_weighted(N, C, edges, g, b)
        if (N, C, P) == (20, 12, 5):
            assert list(qmap) == expect
        # every logical qubit appears once
        assert sorted(qmap) == list(range(N))


@pytest.mark.parametrize("case", [0, 1, 2])
def test_run_vs_kn_closed_form(case):
    c = GU.kn_cases()[case]
    N = c["N"]
    edges = G.complete(N)
    sv, qmap, W = O.run_weighted(N, 5, edges, c["gammas"], c["betas"])
    pc = GU.popcount(np.arange(1 << N))
    np.testing.assert_allclose(sv, c["phi"][pc], rtol=0, atol=1e-13 * 2.0 ** (-N / 2) * 1e3)
    assert O.expect_cut(sv, W) == pytest.approx(c["expect_cut"], rel=1e-12)


@pytest.mark.parametrize("N,seed", [(10, 31), (12, 32), (14, 33)])
def test_run_p1_analytic(N, seed):
    edges = G.gnp(N, 0.5, seed, "uniform")
    g, b = G.angles(1, seed)
    sv, qmap, W = O.run_weighted(N, 4, edges, g, b)
    assert O.expect_cut(sv, W) == pytest.approx(R.p1_expect_cut(N, edges, g[0], b[0]), rel=1e-12)


def test_p1_analytic_fixture_graph_n20():
    d = GU.instances()["n20_gnp05"]
    p1 = GU.p1_analytic()["n20_gnp05"]
    sv, qmap, W = O.run_weighted(20, 12, d["edges"], [p1["gamma"]], [p1["beta"]])
    assert O.expect_cut(sv, W) == pytest.approx(p1["expect_cut"], rel=1e-12)


def test_mixer_on_plus_state_closed_form():
    # gamma = 0: RX(beta)|+> = exp(-i beta/2)|+>, so psi = exp(-i N beta P/2) 2^(-N/2)
    N, P, beta = 9, 3, 0.7
    edges = G.gnp(N, 0.5, 4, "uniform")
    sv, _, _ = O.run_weighted(N, 4, edges, [0.0] * P, [beta] * P)
    np.testing.assert_allclose(sv, np.exp(-1j * N * beta * P / 2) * 2.0 ** (-N / 2), atol=1e-15)


def test_norm_topk_logical_helpers():
    N = 10
    rng = np.random.default_rng(3)
    sv = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    assert O.norm2(sv) == pytest.approx(float(np.sum(np.abs(sv) ** 2)), rel=1e-13)
    idx, prob = O.topk(sv, 8)
    p = np.abs(sv) ** 2
    order = np.lexsort((np.arange(1 << N), -p))[:8]
    np.testing.assert_array_equal(idx, order)
    np.testing.assert_allclose(prob, p[order], rtol=1e-15)
    qmap = np.array([3, 0, 9, 1, 2, 8, 4, 5, 7, 6], dtype=np.int32)
    np.testing.assert_array_equal(O.to_logical(sv, qmap), R.logical_from_physical(sv, qmap))


def test_cut_value_vs_zstrings():
    N = 8
    edges = G.gnp(N, 0.5, 9, "uniform")
    W = O.dense_w(N, edges)
    d = R.dense_cut_diag(N, edges)
    for b in range(1 << N):
        assert O.cut_value(W, b) == pytest.approx(d[b], abs=1e-13)


def test_fixture_instances_are_reproducible():
    inst = GU.instances()
    for name in ["n20_gnp05", "n30_reg3", "n30_gnp05", "n34_gnp03_int", "n26_reg4"]:
        w = G.workload(name)
        assert [tuple(e) for e in inst[name]["edges"]] == [tuple(e) for e in w["edges"]]
        np.testing.assert_array_equal(inst[name]["gammas"], w["gammas"])


def test_oracle_n20_regression_fixture():
    d = GU.instances()["n20_gnp05"]
    fx = GU.oracle_n20()
    sv, qmap, W = O.run_weighted(20, 12, d["edges"], d["gammas"], d["betas"])
    assert list(qmap) == fx["qmap"]
    assert O.expect_cut(sv, W) == fx["expect_cut"]
    s = np.array(fx["sample_idx"])
    np.testing.assert_array_equal(sv[s].real, fx["sample_re"])
    np.testing.assert_array_equal(sv[s].imag, fx["sample_im"])


def test_reference_recipe_is_permutation_symmetric():
    # validate_state.py's recipe (K_N, gamma = beta = pi/4) cannot see layout
    # errors: the state depends only on the Hamming weight
    N = 10
    sv, qmap, _ = O.run_weighted(N, 3, G.complete(N), [math.pi / 4] * 2, [math.pi / 4] * 2)
    pc = GU.popcount(np.arange(1 << N))
    for h in range(N + 1):
        vals = sv[pc == h]
        assert np.max(np.abs(vals - vals[0])) < 1e-15


@pytest.mark.parametrize("N,C,P,seed", [(9, 4, 3, 41), (12, 5, 4, 42), (14, 13, 2, 43)])
def test_final_state_is_complement_symmetric(N, C, P, seed):
    # psi(b) = psi(~b) for Max-Cut QAOA from |+>: Cut(b) = Cut(~b), the mixer
    # commutes with X^{(x)N} and |+> is X^{(x)N}-invariant.  The half-state
    # schedule stores z_{N-1} = 0 only and rests on this; the dense
    # textbook-gate simulator shows it independently of the oracle.
    edges = G.gnp(N, 0.5, seed, "uniform")
    g, b = G.angles(P, seed)
    sv, qmap, W = O.run_weighted(N, C, edges, g, b)
    lg = O.to_logical(sv, qmap)
    full = (1 << N) - 1
    idx = np.arange(1 << N)
    np.testing.assert_allclose(lg, lg[full ^ idx], rtol=0, atol=1e-15)
    if N <= 10:
        psi = R.dense_qaoa(N, edges, g, b)
        np.testing.assert_allclose(psi, psi[full ^ idx], rtol=0, atol=1e-14)


@pytest.mark.parametrize("name", ["n30_gnp05", "n30_reg3"])
def test_large_fixtures_are_qaoa_states(name):
    # the N=30 P=5 oracle fixtures (tests/golden/make_oracle_large.py): unit
    # norm, <C> inside [0, W], and the complement symmetry psi(b) = psi(~b)
    # of every QAOA state seen in the logical top-16: each bitstring's
    # complement is there with the same probability (a wrong qubit map or
    # logical index would break the pairing)
    import os
    d = GU.instances()[name]
    with np.load(os.path.join(GU.GOLDEN, f"oracle_{name}.npz"), allow_pickle=False) as z:
        fx = {k: z[k] for k in z.files}
    N = d["N"]
    assert abs(float(fx["norm2"]) - 1.0) < 1e-13
    assert 0.0 < float(fx["expect_cut"]) < sum(w for _, _, w in d["edges"])
    full = (1 << N) - 1
    prob = {int(i): float(p) for i, p in zip(fx["top16_idx"], fx["top16_prob"])}
    paired = [i for i in prob if (full ^ i) in prob]
    assert len(paired) >= 14
    for i in paired:
        assert prob[i] == pytest.approx(prob[full ^ i], rel=1e-12)
    assert len(fx["sample_idx"]) > 8000
    amps = fx["sample_re"] + 1j * fx["sample_im"]
    # sampled |psi|^2 average ~ 2^-N
    assert np.mean(np.abs(amps) ** 2) * (1 << N) == pytest.approx(1.0, rel=0.2)


def test_n26_fixture_expectations_in_range():
    import os
    d = GU.instances()["n26_reg4"]
    with np.load(os.path.join(GU.GOLDEN, "oracle_n26_reg4.npz"), allow_pickle=False) as z:
        ex = z["expect_cut"]
        assert z["gammas"].shape == (16, d["P"])
    W = sum(w for _, _, w in d["edges"])
    assert np.all((ex > 0) & (ex < W))
