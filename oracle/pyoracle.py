"""ctypes loader for the CPU oracle (oracle/build/libfq_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never loads
it.  Each wrapper names the reference lines its C function restates (paths
relative to the FOR-QAOA repository root).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libfq_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        up = ctypes.POINTER(ctypes.c_uint64)
        vp = ctypes.c_void_p
        L.oq_init_amp.argtypes = [ctypes.c_int]
        L.oq_init_amp.restype = ctypes.c_double
        L.oq_cost_layer.argtypes = [ctypes.c_int, vp, dp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.oq_multi_rx.argtypes = [ctypes.c_int, ctypes.c_int, vp, ip, ctypes.c_int, ctypes.c_double]
        L.oq_multi_rx.restype = ctypes.c_int
        L.oq_swap_conti.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oq_swap_conti.restype = ctypes.c_int
        L.oq_relabel.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oq_relabel.restype = ctypes.c_int
        L.oq_run_weighted.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp, vp, ip]
        L.oq_run_weighted.restype = ctypes.c_int
        L.oq_run_weighted_timed.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp, vp, ip, dp]
        L.oq_run_weighted_timed.restype = ctypes.c_int
        L.oq_cut_value.argtypes = [ctypes.c_int, dp, ctypes.c_uint64]
        L.oq_cut_value.restype = ctypes.c_double
        L.oq_expect_cut.argtypes = [ctypes.c_int, vp, dp]
        L.oq_expect_cut.restype = ctypes.c_double
        L.oq_norm2.argtypes = [ctypes.c_int, vp]
        L.oq_norm2.restype = ctypes.c_double
        L.oq_topk.argtypes = [ctypes.c_int, vp, ctypes.c_int, up, dp]
        L.oq_to_logical.argtypes = [ctypes.c_int, vp, ip, vp]
        L.oq_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def dense_w(N, edges):
    """Upper-triangular N*N weight matrix W[min][max] as addEdgeTo2DGraph
    (src/graph.cpp:54-64) builds it: a repeated edge overwrites."""
    W = np.zeros((N, N), dtype=np.float64)
    for u, v, w in edges:
        a, b = (u, v) if u < v else (v, u)
        W[a, b] = w
    return W


def init_amp(N):
    """initH (src/rzz.cu:26-28)."""
    return lib().oq_init_amp(N)


def cost_layer(sv, W, gamma, first):
    """_rotationCompressionWeighted_conti (src/rzz.cu:121-156)."""
    N = W.shape[0]
    W = np.ascontiguousarray(W, dtype=np.float64)
    lib().oq_cost_layer(N, sv.ctypes.data, _dp(W), gamma, int(first), init_amp(N))


def multi_rx(sv, N, C, targets, beta):
    """multiRotateX (src/rx.cu:118-174)."""
    t = np.ascontiguousarray(targets, dtype=np.int32)
    rc = lib().oq_multi_rx(N, C, sv.ctypes.data, _ip(t), len(t), beta)
    if rc:
        raise ValueError("oq_multi_rx: bad targets")


def swap_conti(sv, N, out, inn, k):
    """swap_gate_conti local branch (src/swap.cu:166-191)."""
    rc = lib().oq_swap_conti(N, sv.ctypes.data, out, inn, k)
    if rc:
        raise ValueError("oq_swap_conti: bad blocks")


def relabel(W, out, inn, k):
    """mapping_after_swap_conti (src/swap.cu:27-50), in place."""
    assert W.flags.c_contiguous and W.dtype == np.float64
    rc = lib().oq_relabel(W.shape[0], _dp(W), out, inn, k)
    if rc:
        raise ValueError("oq_relabel: bad blocks")


def run_weighted(N, C, edges, gammas, betas):
    """Driver layer loop (src/Weighted.cpp:93-137) with per-layer angles.

    Returns (state in physical order, qmap phys->logical, relabelled W)."""
    P = len(gammas)
    assert len(betas) == P
    W = dense_w(N, edges)
    g = np.ascontiguousarray(gammas, dtype=np.float64)
    b = np.ascontiguousarray(betas, dtype=np.float64)
    sv = np.zeros(1 << N, dtype=np.complex128)
    qmap = np.zeros(N, dtype=np.int32)
    rc = lib().oq_run_weighted(N, C, P, _dp(W), _dp(g), _dp(b), sv.ctypes.data, _ip(qmap))
    if rc:
        raise ValueError(f"oq_run_weighted failed ({rc})")
    return sv, qmap, W


def run_weighted_timed(N, C, edges, gammas, betas):
    """run_weighted, also returning the seconds of the P-layer loop alone
    (where MEASURET_START / MEASURET_END sit, src/Weighted.cpp:113,140)."""
    P = len(gammas)
    assert len(betas) == P
    W = dense_w(N, edges)
    g = np.ascontiguousarray(gammas, dtype=np.float64)
    b = np.ascontiguousarray(betas, dtype=np.float64)
    sv = np.zeros(1 << N, dtype=np.complex128)
    qmap = np.zeros(N, dtype=np.int32)
    secs = np.zeros(1, dtype=np.float64)
    rc = lib().oq_run_weighted_timed(N, C, P, _dp(W), _dp(g), _dp(b), sv.ctypes.data, _ip(qmap), _dp(secs))
    if rc:
        raise ValueError(f"oq_run_weighted_timed failed ({rc})")
    return sv, qmap, W, float(secs[0])


def to_logical(sv, qmap):
    N = len(qmap)
    q = np.ascontiguousarray(qmap, dtype=np.int32)
    out = np.empty_like(sv)
    lib().oq_to_logical(N, sv.ctypes.data, _ip(q), out.ctypes.data)
    return out


def expect_cut(sv, W):
    N = W.shape[0]
    W = np.ascontiguousarray(W, dtype=np.float64)
    return lib().oq_expect_cut(N, sv.ctypes.data, _dp(W))


def cut_value(W, b):
    W = np.ascontiguousarray(W, dtype=np.float64)
    return lib().oq_cut_value(W.shape[0], _dp(W), int(b))


def norm2(sv):
    N = int(sv.size).bit_length() - 1
    return lib().oq_norm2(N, sv.ctypes.data)


def topk(sv, k):
    N = int(sv.size).bit_length() - 1
    idx = np.zeros(k, dtype=np.uint64)
    prob = np.zeros(k, dtype=np.float64)
    lib().oq_topk(N, sv.ctypes.data, k, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), _dp(prob))
    return idx, prob


def num_threads():
    return lib().oq_num_threads()
