"""fq_expect_cut_static (allocation-free <C>) against fq_expect_cut.

Both run the same reduction kernels; the static entry point keeps their
scratch in the library's fixed device array.  Two calls in a row, the second
on a smaller state of the other amplitude type, each equal to fq_expect_cut
on the same state (the same kernels in the same order: equal to the last
bit; the bar is 1e-15 relative); and a state beyond the array's limit
(N > 30) is rejected with FQ_ERR_ARG before any device work.
"""
import numpy as np
import pytest

from forqaoa import FQError, Graph, State, _lib
from forqaoa import graphs as G

pytestmark = pytest.mark.gpu


def _prepared(N, dtype, seed):
    edges = G.gnp(N, 0.5, seed, "uniform")
    g, b = G.angles(3, seed)
    st = State(N, C=min(N, 13 if dtype == "c128" else 14), dtype=dtype)
    gr = Graph(N, edges)
    st.run(gr, g, b, "fused")
    return st, gr


@pytest.mark.parametrize("first,second", [((22, "c128"), (14, "c64")), ((20, "c64"), (9, "c128"))])
def test_static_expect_matches_allocating_entry_point(first, second):
    for (N, dtype), seed in ((first, 51), (second, 52)):
        print(f"static <C>: N={N} {dtype}", flush=True)
        st, gr = _prepared(N, dtype, seed)
        try:
            e_ref, n_ref = st.expect_cut(gr)
            e_st, n_st = st.expect_cut_static(gr)
            assert e_st == pytest.approx(e_ref, rel=1e-15, abs=0)
            assert n_st == pytest.approx(n_ref, rel=1e-15, abs=0)
            # a second call on the same state reuses the array (nothing is cleared)
            e2, _ = st.expect_cut_static(gr)
            assert e2 == e_st
        finally:
            st.destroy()
            gr.destroy()


def test_static_expect_rejects_states_beyond_the_limit():
    st = State(31, C=14, dtype="c64")
    gr = Graph(31, G.gnp(31, 0.2, 53, "uniform"))
    try:
        with pytest.raises(FQError) as e:
            st.expect_cut_static(gr)
        assert e.value.code == _lib.FQ_ERR_ARG
        assert "N <= 30" in str(e.value)
    finally:
        st.destroy()
        gr.destroy()
