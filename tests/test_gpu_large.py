"""Full-size runs: N=34 complex64 (137 GB) and N=32 complex128 vs complex64.

The oracle cannot run at these sizes; size-independent pins instead:
* K_34 with unit weights: sampled amplitudes and <C> against the
  Hamming-weight closed form (tests/golden/kn_closed_form.json);
* G(34, 0.3) with integer weights, one layer: <C> against the analytic
  single-layer expectation;
* N=32, P=10, G(32, 0.3) integer weights: the complex64 run against the
  complex128 run of the same instance (the error of the complex64 path is
  printed and bounded);
* N=34, P=10, the headline complex64 instance, against complex128: a full
  complex128 state of 34 qubits (275 GB) does not fit beside the rest, but
  the half state z_33 = 0 does -- a one-slot complex128 half-state slab
  (2^33 x 16 B = 137 GB), whose <C> is the state's (psi(b) = psi(~b));
* norm preservation.
"""
import numpy as np
import pytest

import golden_util as GU
import refmath as R
from forqaoa import Graph, State
from forqaoa import graphs as G
from forqaoa import _lib
from forqaoa.serve import Slab

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("schedule", ["fused", "z2"])
def test_n34_c64_complete_graph_closed_form(schedule):
    # (both schedules run the complex64 layer-boundary pass whose lane bits 4
    # and 5 get their RX through lane swaps)
    c = [k for k in GU.kn_cases() if k["N"] == 34][0]
    N = 34
    st = State(N, C=14, dtype="c64")
    g = Graph(N, G.complete(N))
    st.run(g, c["gammas"], c["betas"], schedule)
    rng = np.random.default_rng(34)
    idx = rng.integers(0, 1 << N, size=8192, dtype=np.uint64)
    got = st.gather(idx, logical=True)
    want = c["phi"][GU.popcount(idx)]
    rms = 2.0 ** (-N / 2)
    err = float(np.max(np.abs(got - want)) / rms)
    ex, n2 = st.expect_cut(g)
    rel = abs(ex - c["expect_cut"]) / c["expect_cut"]
    print(f"N=34 c64 K_34 P=2 {schedule}: max amplitude error / 2^-17 = {err:.2e}, <C> rel err {rel:.2e}, "
          f"norm2-1 = {n2 - 1:.2e}", flush=True)
    assert err < 1e-4
    assert rel < 1e-5
    assert abs(n2 - 1) < 1e-5
    st.destroy()


def test_n34_c64_single_layer_analytic():
    w = G.workload("n34_gnp03_int")
    N = 34
    st = State(N, C=14, dtype="c64")
    g = Graph(N, w["edges"])
    st.run(g, w["gammas"][:1], w["betas"][:1], "fused")
    ex, n2 = st.expect_cut(g)
    ref = R.p1_expect_cut(N, w["edges"], w["gammas"][0], w["betas"][0])
    rel = abs(ex - ref) / abs(ref)
    print(f"N=34 c64 G(34,0.3) int weights P=1: <C> {ex:.10g} vs analytic {ref:.10g} (rel {rel:.2e})",
          flush=True)
    assert rel < 1e-5
    st.destroy()


def test_n34_c64_p10_runs_and_preserves_norm():
    w = G.workload("n34_gnp03_int")
    N = 34
    st = State(N, C=14, dtype="c64")
    g = Graph(N, w["edges"])
    st.run(g, w["gammas"], w["betas"], "fused")
    ex, n2 = st.expect_cut(g)
    print(f"N=34 c64 P=10: <C> = {ex:.10g}, norm2 - 1 = {n2 - 1:.2e}", flush=True)
    assert abs(n2 - 1) < 1e-4
    assert 0 < ex < sum(x[2] for x in w["edges"])
    st.destroy()


def test_n32_c64_error_vs_c128():
    w = G.workload("n32_gnp03_int")
    N, P = 32, w["P"]
    g = Graph(N, w["edges"])
    rng = np.random.default_rng(32)
    idx = rng.integers(0, 1 << N, size=8192, dtype=np.uint64)
    a = State(N, C=13, dtype="c128")
    a.run(g, w["gammas"], w["betas"], "fused")
    e128, n128 = a.expect_cut(g)
    v128 = a.gather(idx, logical=True)
    a.destroy()
    b = State(N, C=14, dtype="c64")
    b.run(g, w["gammas"], w["betas"], "fused")
    e64, n64 = b.expect_cut(g)
    v64 = b.gather(idx, logical=True)
    b.destroy()
    rms = 2.0 ** (-N / 2)
    amp_err = float(np.max(np.abs(v64 - v128)) / rms)
    rel = abs(e64 - e128) / abs(e128)
    print(f"N=32 P=10 G(32,0.3) int: complex64 vs complex128 <C> rel err {rel:.2e}, "
          f"max sampled amplitude error / 2^-16 = {amp_err:.2e}, norm2 c128-1 {n128 - 1:.2e}", flush=True)
    assert rel < 1e-5
    assert amp_err < 1e-3
    assert abs(n128 - 1) < 1e-11


def test_n34_c64_p10_vs_c128_half_state():
    w = G.workload("n34_gnp03_int")
    N = 34
    _lib.check(_lib.lib().fq_set_memory_budget(0))   # no budget (another test may have left one)
    g = Graph(N, w["edges"])
    try:
        st = State(N, C=14, dtype="c64")
        try:
            st.run(g, w["gammas"], w["betas"], "fused")
            e64, n64 = st.expect_cut(g)
        finally:
            st.destroy()
        print(f"N=34 c64 P=10 fused: <C> {e64:.12g}, norm2-1 {n64 - 1:.2e}", flush=True)
        sb = Slab(N, 1, "c128", "z2")
        try:
            sb.assign(0, g, type("R", (), {"gammas": w["gammas"], "betas": w["betas"]})())
            out = []
            while not out:
                sb.step()
                out = sb.collect()
            e128, n128 = out[0].expect, out[0].norm2
        finally:
            sb.destroy()
    finally:
        g.destroy()
    rel = abs(e64 - e128) / abs(e128)
    print(f"N=34 P=10 G(34,0.3) int: complex128 half-state <C> {e128:.12g} (norm2-1 {n128 - 1:.2e}); "
          f"complex64 rel err {rel:.2e}", flush=True)
    assert abs(n128 - 1) < 1e-11
    assert rel < 1e-5
