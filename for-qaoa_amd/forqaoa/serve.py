"""Continuous-batching QAOA evaluation server over packed slabs.

A request is one QAOA evaluation: a graph (resident by key, or new), depth P
and per-layer angles; the answer is <C>.  Per qubit count N the server keeps
one slab (fq_slab_*): 2^s state slots of 2^N amplitudes packed so that each
tile pass is a single launch over every active slot.  The scheduler loop

  * admits arrived requests into free slots of their N (the slab is created on
    first use, under the process-wide HBM budget fq_set_memory_budget; a
    request whose slab cannot be created waits until memory is released),
  * advances every slab whose previous step finished by one fused layer
    (fq_slab_step, asynchronous on the slab's own stream, so slabs of
    different N overlap on the device),
  * collects finished slots (fq_slab_collect) and frees them for the next
    request -- requests join and leave at layer boundaries while others keep
    running.

Graphs stay resident (host weights and their device cost plans) across
requests: re-evaluating a graph with new angles, as an outer optimiser does,
costs no graph upload.

The reference has no serving path; this is the serving form named for this
build.  `one_at_a_time` is the plain baseline: for each request in turn a
fresh state and graph, the reference layer schedule (fq_run REFERENCE), <C>,
release.
"""
from __future__ import annotations

import ctypes
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import FQ_C64, FQ_C128, DTYPES, SCHEDULES, FQError, Graph, State, _lib
from ._lib import check, lib


@dataclass
class Request:
    rid: int
    N: int
    P: int
    graph_key: int
    gammas: np.ndarray
    betas: np.ndarray
    arrival: float = 0.0
    # filled by the server
    start: float = float("nan")
    done: float = float("nan")
    expect: float = float("nan")
    norm2: float = float("nan")
    joined_running: int = 0


@dataclass
class Stream:
    """A synthetic request stream: graphs by key and the requests."""
    graphs: dict = field(default_factory=dict)   # key -> (N, edges)
    requests: list = field(default_factory=list)


def make_stream(n_requests, seed=20251207, n_range=(16, 30), p_range=(1, 10), gnp_p=(0.2, 0.8),
                reuse=0.8, rate=None):
    """Requests with N uniform in n_range, P uniform in p_range, graphs G(N, p)
    with p uniform in gnp_p and weights uniform in (0, 1], random angles; a
    fraction `reuse` re-evaluate an existing graph with new angles.  Arrival
    times are Poisson with the given rate (requests/s), or all 0."""
    from . import graphs as G

    rng = np.random.default_rng(seed)
    st = Stream()
    t = 0.0
    for rid in range(n_requests):
        if st.graphs and rng.random() < reuse:
            key = int(rng.integers(0, len(st.graphs)))
            N = st.graphs[key][0]
        else:
            key = len(st.graphs)
            N = int(rng.integers(n_range[0], n_range[1] + 1))
            p = float(rng.uniform(*gnp_p))
            st.graphs[key] = (N, G.gnp(N, p, int(rng.integers(0, 2**31)), "uniform"))
        P = int(rng.integers(p_range[0], p_range[1] + 1))
        g = rng.uniform(0, math.pi, P)
        b = rng.uniform(0, math.pi, P)
        if rate:
            t += float(rng.exponential(1.0 / rate))
        st.requests.append(Request(rid, N, P, key, g, b, arrival=t))
    return st


def in_use():
    """HBM held by this process's states and slabs (fq_memory_in_use)."""
    used = ctypes.c_size_t()
    check(lib().fq_memory_in_use(ctypes.byref(used)))
    return used.value


def default_slots(N):
    """Slots per slab: about 2^24 amplitudes per slab for small N (packing),
    two slots for large N."""
    return int(max(2, min(64, 2 ** max(0, 24 - N))))


class Slab:
    """K packed slots of one N; schedule "fused" (full states) or "z2" (each
    slot stores the half z_{N-1} = 0 of its state, psi(b) = psi(~b))."""

    def __init__(self, N, nslots, dtype, schedule="fused"):
        h = ctypes.c_void_p()
        check(lib().fq_slab_create_sched(int(N), int(nslots), DTYPES[dtype], SCHEDULES[schedule],
                                         ctypes.byref(h)))
        self.h = h
        self.schedule = schedule
        self.N = N
        self.K = nslots
        self.free = list(range(nslots))
        self.slot_req = {}
        self.busy = False
        self._buf_i = (ctypes.c_int * nslots)()
        self._buf_e = (ctypes.c_double * nslots)()
        self._buf_n = (ctypes.c_double * nslots)()

    def assign(self, slot, graph, req):
        g = np.ascontiguousarray(req.gammas, dtype=np.float64)
        b = np.ascontiguousarray(req.betas, dtype=np.float64)
        check(lib().fq_slab_assign(self.h, slot, graph.handle, len(g),
                                   g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   b.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        self.slot_req[slot] = req

    def set_layers_per_step(self, layers):
        """Layers a step runs per slot (slabs of large slots; see
        fq_slab_set_layers_per_step)."""
        check(lib().fq_slab_set_layers_per_step(self.h, int(layers)))

    def step(self):
        n = ctypes.c_int()
        check(lib().fq_slab_step(self.h, ctypes.byref(n)))
        self.busy = bool(self.slot_req)
        return n.value

    def poll(self):
        b = ctypes.c_int()
        check(lib().fq_slab_poll(self.h, ctypes.byref(b)))
        self.busy = bool(b.value)
        return self.busy

    def collect(self):
        n = ctypes.c_int()
        check(lib().fq_slab_collect(self.h, self._buf_i, self._buf_e, self._buf_n, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            slot = self._buf_i[i]
            req = self.slot_req.pop(slot)
            req.expect = self._buf_e[i]
            req.norm2 = self._buf_n[i]
            self.free.append(slot)
            out.append(req)
        return out

    def copy_out(self, slot, dtype):
        npdt = np.complex128 if DTYPES[dtype] == FQ_C128 else np.complex64
        a = np.empty(1 << self.N, dtype=npdt)
        check(lib().fq_slab_copy_out(self.h, slot, a.ctypes.data, a.nbytes))
        return a

    def destroy(self):
        if self.h is not None:
            lib().fq_slab_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Server:
    """schedule: "z2" (half-state slabs for N >= 3, the default) or "fused"
    (full-state slabs)."""

    def __init__(self, dtype="c64", budget_bytes=200e9, slots_for=default_slots, schedule="z2"):
        self.dtype = dtype
        self.slots_for = slots_for
        self.schedule = schedule
        self.budget_bytes = budget_bytes
        check(lib().fq_set_memory_budget(ctypes.c_size_t(int(budget_bytes))))
        self.slabs = {}
        self.graphs = {}
        self.peak_in_use = in_use()  # HBM held by this process, sampled after every slab creation

    def graph(self, stream, key):
        g = self.graphs.get(key)
        if g is None:
            N, edges = stream.graphs[key]
            g = Graph(N, edges)
            self.graphs[key] = g
        return g

    def _slab(self, N, waiting):
        sb = self.slabs.get(N)
        while sb is None:
            try:
                sched = self.schedule if N >= 3 else "fused"
                sb = Slab(N, self.slots_for(N), self.dtype, sched)
            except FQError as e:
                if e.code != _lib.FQ_ERR_BUDGET:
                    raise
                # release an idle slab of another size, if any
                idle = [M for M, s in self.slabs.items()
                        if not s.slot_req and not s.busy and not waiting.get(M)]
                if not idle:
                    return None
                self.slabs.pop(idle[0]).destroy()
                continue
            self.slabs[N] = sb
            self.peak_in_use = max(self.peak_in_use, in_use())
        return sb

    def run(self, stream, realtime=True, log=None):
        """Serve every request of the stream; returns the requests with
        start/done times (seconds from the start) and <C>."""
        reqs = sorted(stream.requests, key=lambda r: r.arrival)
        incoming = deque(reqs)
        waiting = {}
        done = []
        t0 = time.perf_counter()
        n_total = len(reqs)
        while len(done) < n_total:
            now = time.perf_counter() - t0
            while incoming and (not realtime or incoming[0].arrival <= now):
                r = incoming.popleft()
                waiting.setdefault(r.N, deque()).append(r)
            progressed = False
            for N in sorted(set(waiting) | set(self.slabs)):
                q = waiting.get(N)
                sb = self.slabs.get(N)
                if sb is None:
                    if not q:
                        continue
                    sb = self._slab(N, waiting)
                    if sb is None:
                        continue
                if sb.busy and sb.poll():
                    continue
                fin = sb.collect() if sb.slot_req else []
                tnow = time.perf_counter() - t0
                for r in fin:
                    r.done = tnow
                    done.append(r)
                while q and sb.free:
                    r = q.popleft()
                    r.joined_running = len(sb.slot_req)  # slots already in flight
                    sb.assign(sb.free.pop(0), self.graph(stream, r.graph_key), r)
                    r.start = time.perf_counter() - t0
                if sb.slot_req:
                    sb.step()
                    progressed = True
            idle = not progressed and all(not s.busy and not s.slot_req for s in self.slabs.values())
            if idle and any(waiting.values()):
                # nothing runs and nothing can be released: these requests
                # can never be admitted under the budget
                blocked = sorted(N for N, q in waiting.items() if q)
                raise FQError(_lib.FQ_ERR_BUDGET,
                              f"requests of N={blocked} do not fit the memory budget "
                              f"({self.budget_bytes:.3g} bytes, {in_use():.3g} in use)")
            if not progressed and incoming and realtime:
                wait = incoming[0].arrival - (time.perf_counter() - t0)
                if wait > 0 and all(not s.busy for s in self.slabs.values()):
                    time.sleep(min(wait, 0.001))
        return done

    def close(self):
        for sb in self.slabs.values():
            sb.destroy()
        self.slabs.clear()
        for g in self.graphs.values():
            g.destroy()
        self.graphs.clear()


def one_at_a_time(stream, requests, dtype="c64", schedule="reference", C=None):
    """The plain baseline: each request alone, start to finish: fresh state and
    graph, the layer schedule, <C>, release.  Returns seconds."""
    t0 = time.perf_counter()
    for r in requests:
        N, edges = stream.graphs[r.graph_key]
        c = C if C is not None else min(N, 13 if DTYPES[dtype] == FQ_C128 else 14)
        st = State(N, C=c, dtype=dtype)
        g = Graph(N, edges)
        st.run(g, r.gammas, r.betas, schedule)
        r.expect, r.norm2 = st.expect_cut(g)
        st.destroy()
        g.destroy()
    return time.perf_counter() - t0


def latency_stats(done):
    """p50 / p99 / mean latency (arrival to result) in ms, overall and per
    qubit-count band (N 16-20, 21-25, 26-28, 29-30), with the time from
    arrival to admission (queueing) beside it."""
    lat = np.array([r.done - r.arrival for r in done])
    wait = np.array([r.start - r.arrival for r in done])
    out = dict(p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
               mean_ms=float(lat.mean() * 1e3), queue_p99_ms=float(np.percentile(wait, 99) * 1e3))
    bands = {}
    for lo, hi in ((0, 20), (21, 25), (26, 28), (29, 40)):
        sel = np.array([lo <= r.N <= hi for r in done])
        if sel.any():
            bands[f"N{max(lo, min(r.N for r in done))}-{min(hi, max(r.N for r in done))}"] = dict(
                n=int(sel.sum()), p50_ms=float(np.percentile(lat[sel], 50) * 1e3),
                p99_ms=float(np.percentile(lat[sel], 99) * 1e3),
                queue_p99_ms=float(np.percentile(wait[sel], 99) * 1e3))
    out["by_n"] = bands
    return out


def bytes_floor(stream, requests, dtype="c64", schedule="fused"):
    """Algorithmic HBM bytes of the slab schedule for these requests: the
    tile passes each request's slot runs (layer 0's first pass writes the
    stored state once, every other pass reads and writes it) plus one <C>
    read.  A half-state slab stores and moves half of each state.  Slots of
    at least 2^10 first-pass tiles run fq_run's pipelined layers (1 + P(m-1)
    passes for m passes per layer, when m >= 2 (full) or 3 (half state));
    smaller slots run P*m packed passes."""
    from . import layer_passes

    elem = 16 if DTYPES[dtype] == FQ_C128 else 8
    qmax = 13 if DTYPES[dtype] == FQ_C128 else 14
    m_of = {}
    tot = 0.0
    for r in requests:
        half = schedule == "z2" and r.N >= 3
        ns = r.N - 1 if half else r.N
        S = (1 << ns) * elem
        if r.N not in m_of:
            m_of[r.N] = len(layer_passes(r.N, dtype, "z2" if half else "fused"))
        m = m_of[r.N]
        per_slot = ns - min(qmax, ns) >= 10
        piped = per_slot and m >= (3 if half else 2)
        npass = 1 + r.P * (m - 1) if piped else r.P * m
        tot += S * (2 * npass - 1) + S
    return tot
