"""forqaoa -- MI355X-native QAOA Max-Cut state-vector simulation.

Python face of libforqaoa (include/forqaoa.h).  The classes mirror the
reference's module boundaries and entry-point names:

=====================================  ==========================================
reference (FOR-QAOA)                   here
=====================================  ==========================================
createMPIState(N, D, B, C)             State(N, D=0, B=..., C=..., dtype=...)
  (src/state.cpp:31-86)
initH (src/rzz.cu:26-28)               State.initH()
init2DGraph / addEdgeTo2DGraph         Graph(N) / Graph.addEdgeTo2DGraph(u, v, w)
  (src/graph.cpp:46-64)
rotationCompressionWeighted            State.rotationCompressionWeighted(angle,
  (src/rzz.cu:173-206)                    graph, isFirstLayer)
multiRotateX (src/rx.cu:118-174)       State.multiRotateX(targets, gate_size, angle)
swap_gate_conti (src/swap.cu:104-202)  State.swap_gate_conti(out, in, num)
mapping_after_swap_conti               Graph.mapping_after_swap_conti(out, in, num)
  (src/swap.cu:27-50)
layer loop (src/Weighted.cpp:97-137)   State.run(graph, gammas, betas, schedule)
=====================================  ==========================================

Errors raise FQError (the reference exits the process instead).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import (
    FQ_C64,
    FQ_C128,
    FQ_K_ADJOINT,
    FQ_K_COST,
    FQ_K_FUSED,
    FQ_K_REDUCE,
    FQ_K_RX,
    FQ_K_SWAP,
    FQ_K_TOPK,
    FQ_SCHED_FUSED,
    FQ_SCHED_REFERENCE,
    FQ_SCHED_Z2,
    FQError,
    KERNEL_CLASSES,
    check,
    lib,
)

DTYPES = {"c128": FQ_C128, "c64": FQ_C64, FQ_C128: FQ_C128, FQ_C64: FQ_C64}
SCHEDULES = {"reference": FQ_SCHED_REFERENCE, "fused": FQ_SCHED_FUSED, "z2": FQ_SCHED_Z2,
             FQ_SCHED_REFERENCE: FQ_SCHED_REFERENCE, FQ_SCHED_FUSED: FQ_SCHED_FUSED,
             FQ_SCHED_Z2: FQ_SCHED_Z2}


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Graph:
    """Dense upper-triangular weighted graph (init2DGraph, src/graph.cpp:46-52)."""

    def __init__(self, N, edges=None):
        h = ctypes.c_void_p()
        check(lib().fq_graph_create(int(N), ctypes.byref(h)))
        self._h = h
        self.N = int(N)
        if edges is not None:
            for u, v, w in edges:
                self.add_edge(u, v, w)

    @property
    def handle(self):
        if self._h is None:
            raise FQError(_lib.FQ_ERR_ARG, "graph already destroyed")
        return self._h

    def add_edge(self, u, v, w):
        check(lib().fq_graph_add_edge(self.handle, int(u), int(v), float(w)))

    addEdgeTo2DGraph = add_edge

    def weights(self):
        W = np.zeros((self.N, self.N), dtype=np.float64)
        check(lib().fq_graph_weights(self.handle, _dptr(W)))
        return W

    def relabel(self, swap_out, swap_in, num_swap):
        check(lib().fq_graph_relabel(self.handle, int(swap_out), int(swap_in), int(num_swap)))

    mapping_after_swap_conti = relabel

    def destroy(self):
        if self._h is not None:
            lib().fq_graph_destroy(self._h)
            self._h = None

    free2DGraph = destroy

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class State:
    """State vector of 2^N amplitudes in HBM (createMPIState + prepare)."""

    def __init__(self, N, C, dtype="c128", D=0, B=None):
        if B is None:
            B = N - D
        h = ctypes.c_void_p()
        check(lib().fq_state_create(int(N), int(D), int(B), int(C), DTYPES[dtype], ctypes.byref(h)))
        self._h = h
        self.N = int(N)
        self.C = int(C)
        self.dtype = DTYPES[dtype]
        self.np_dtype = np.complex128 if self.dtype == FQ_C128 else np.complex64

    @property
    def handle(self):
        if self._h is None:
            raise FQError(_lib.FQ_ERR_ARG, "state already destroyed")
        return self._h

    @property
    def nbytes(self):
        n = ctypes.c_size_t()
        check(lib().fq_state_info(self.handle, None, None, None, ctypes.byref(n)))
        return n.value

    def device_ptr(self):
        p = ctypes.c_void_p()
        check(lib().fq_state_device_ptr(self.handle, ctypes.byref(p)))
        return p.value

    def set_stream(self, stream_ptr):
        check(lib().fq_state_set_stream(self.handle, ctypes.c_void_p(stream_ptr)))

    def init_zero(self):
        check(lib().fq_state_init_zero(self.handle))

    stateInitDeviceState = init_zero

    def init_h(self):
        check(lib().fq_init_h(self.handle))

    initH = init_h

    def cost_layer(self, graph, gamma, first):
        check(lib().fq_cost_layer(self.handle, graph.handle, float(gamma), int(bool(first))))

    def rotationCompressionWeighted(self, angle, graph, isFirstLayer):
        self.cost_layer(graph, angle, isFirstLayer)

    def multi_rx(self, targets, beta):
        t = np.ascontiguousarray(targets, dtype=np.int32)
        check(lib().fq_multi_rx(self.handle, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(t), float(beta)))

    def multiRotateX(self, targetQubits, gate_size, angle):
        self.multi_rx(list(targetQubits)[:gate_size], angle)

    def swap_conti(self, swap_out, swap_in, num_swap):
        check(lib().fq_swap_conti(self.handle, int(swap_out), int(swap_in), int(num_swap)))

    swap_gate_conti = swap_conti

    def run(self, graph, gammas, betas, schedule="fused"):
        g = np.ascontiguousarray(gammas, dtype=np.float64)
        b = np.ascontiguousarray(betas, dtype=np.float64)
        if g.shape != b.shape:
            raise FQError(_lib.FQ_ERR_ARG, "gammas and betas differ in length")
        check(lib().fq_run(self.handle, graph.handle, len(g), _dptr(g), _dptr(b), SCHEDULES[schedule]))

    def backward(self, graph, gammas, betas):
        """Adjoint gradient of <C> after run(graph, gammas, betas, "fused" or "z2").

        Consumes the state.  Returns (<C>, dC/dgammas, dC/dbetas)."""
        g = np.ascontiguousarray(gammas, dtype=np.float64)
        b = np.ascontiguousarray(betas, dtype=np.float64)
        gg = np.zeros(len(g))
        gb = np.zeros(len(b))
        e = ctypes.c_double()
        check(lib().fq_backward(self.handle, graph.handle, len(g), _dptr(g), _dptr(b), ctypes.byref(e),
                                _dptr(gg), _dptr(gb)))
        return e.value, gg, gb

    def expect_cut(self, graph):
        e = ctypes.c_double()
        n = ctypes.c_double()
        check(lib().fq_expect_cut(self.handle, graph.handle, ctypes.byref(e), ctypes.byref(n)))
        return e.value, n.value

    def expect_cut_static(self, graph):
        """expect_cut without device allocation (fq_expect_cut_static: N <= 30,
        scratch in a fixed device array shared by every call on the GPU)."""
        e = ctypes.c_double()
        n = ctypes.c_double()
        check(lib().fq_expect_cut_static(self.handle, graph.handle, ctypes.byref(e), ctypes.byref(n)))
        return e.value, n.value

    def topk(self, k):
        idx = np.zeros(k, dtype=np.uint64)
        prob = np.zeros(k, dtype=np.float64)
        check(lib().fq_topk(self.handle, int(k), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), _dptr(prob)))
        return idx, prob

    def qubit_map(self):
        m = np.zeros(self.N, dtype=np.int32)
        check(lib().fq_qubit_map(self.handle, m.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return m

    def copy_out(self, logical=False):
        out = np.empty(1 << self.N, dtype=self.np_dtype)
        check(lib().fq_state_copy_out(self.handle, out.ctypes.data, out.nbytes, int(bool(logical))))
        return out

    def copy_in(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.np_dtype)
        if a.size != (1 << self.N):
            raise FQError(_lib.FQ_ERR_ARG, "wrong number of amplitudes")
        check(lib().fq_state_copy_in(self.handle, a.ctypes.data, a.nbytes))

    def gather(self, idx, logical=True):
        i = np.ascontiguousarray(idx, dtype=np.uint64)
        out = np.zeros(2 * i.size, dtype=np.float64)
        check(lib().fq_state_gather(self.handle, i.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), i.size,
                                    _dptr(out), int(bool(logical))))
        return out[0::2] + 1j * out[1::2]

    def clone_from(self, other):
        check(lib().fq_state_clone(self.handle, other.handle))

    def sync(self):
        check(lib().fq_sync(self.handle))

    def timing(self, on=True):
        check(lib().fq_timing_enable(self.handle, int(bool(on))))

    def timing_get(self, kind):
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        by = ctypes.c_double()
        check(lib().fq_timing_get(self.handle, int(kind), ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(by)))
        return ms.value, cnt.value, by.value

    def timing_reset(self):
        check(lib().fq_timing_reset(self.handle))

    def destroy(self):
        if self._h is not None:
            lib().fq_state_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def layer_passes(N, dtype="c128", schedule="fused"):
    """The tile passes of one fused or half-state layer: a list of
    (tile positions, RX targets) per pass; pass 0 applies the cost phase."""
    mx = 8
    n, q = ctypes.c_int(), ctypes.c_int()
    tiles = np.full(16 * mx, -1, dtype=np.int32)
    gates = np.zeros(mx, dtype=np.uint64)
    check(lib().fq_layer_passes(int(N), DTYPES[dtype], SCHEDULES[schedule], mx, ctypes.byref(n), ctypes.byref(q),
                                tiles.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                gates.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
    out = []
    for p in range(n.value):
        t = [int(x) for x in tiles[16 * p:16 * p + q.value]]
        g = [i for i in range(N) if (int(gates[p]) >> i) & 1]
        out.append((t, g))
    return out


def stream_probe(nbytes, reps=3):
    """In-place, read-only and write-only streaming rates (GB/s of HBM traffic)
    of simple kernels over an nbytes buffer: what a pass that streams a state
    of that size could reach on this device."""
    ip, rd, wr = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    check(lib().fq_stream_probe(int(nbytes), int(reps), ctypes.byref(ip), ctypes.byref(rd), ctypes.byref(wr)))
    return dict(inplace_gbs=ip.value, read_gbs=rd.value, write_gbs=wr.value, bytes=int(nbytes))


def init_amp(N):
    """(1/sqrt 2)^N as initH computes it (src/rzz.cu:26-28)."""
    return math.pow(1.0 / math.sqrt(2.0), N)
