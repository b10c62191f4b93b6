#!/usr/bin/env python3
"""Benchmark: QAOA layers/s of weighted Max-Cut state-vector simulation on one
MI355X, with the achieved HBM bandwidth of the dominant kernel.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]

Workload (value): N=30, P=5 weighted Max-Cut on the seeded G(30, 0.5) graph
with weights uniform in (0, 1], per-layer angles uniform in [0, pi),
complex128 amplitudes, the fused schedule (cost phase fused into a tile
pass, RX on the high qubits through strided LDS tiles, layers pipelined: the
pass that ends layer p also applies the cost phase and RX of layer p+1 on its
qubits, 11 tile passes for P=5).  One step is one full P=5 evaluation: the
launch-control layer through layer 5, inputs and state resident in HBM.
value = 5 * steps / seconds.

Also reported (extra): the same evaluation on the 3-regular graph, in
complex64, with the reference schedule (cost, RX chunks, transpositions) in
both amplitude types, and
N=34 P=10 complex64 on G(34, 0.3) with integer weights, and the half-state
schedule (layers on the z_{N-1} = 0 half, psi(b) = psi(~b); the full state is
written after the last layer) for N=30 and N=34.

cpu_baseline: the oracle (the C restatement of the reference's layer loop,
complex128, OpenMP, OMP_PROC_BIND=close) on the host cores, run first, in a
child process: the first two full layers of the same N=30 instance timed
around the layer loop alone (value: layers/s), and the whole N=20 P=5 loop.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "for-qaoa_amd"))

HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "fused_pass_n30_c128_traffic.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_eval(st, g, w, schedule):
    st.run(g, w["gammas"], w["betas"], schedule)


def measure(name, dtype, schedule, steps, warmup, C=None, dominant=None):
    from forqaoa import Graph, State, KERNEL_CLASSES
    from forqaoa import graphs as G

    w = G.workload(name)
    N, P = w["N"], w["P"]
    if C is None:
        C = w["C"] if dtype == "c128" else min(14, w["C"] + 1)
    st = State(N, C=C, dtype=dtype)
    g = Graph(N, w["edges"])
    for _ in range(warmup):
        run_eval(st, g, w, schedule)
    st.sync()
    st.timing(True)
    st.timing_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        run_eval(st, g, w, schedule)
    st.sync()
    t1 = time.perf_counter()
    secs = t1 - t0
    kern = {}
    total_bytes = 0.0
    for kind, kname in enumerate(KERNEL_CLASSES):
        ms, cnt, by = st.timing_get(kind)
        if cnt:
            kern[kname] = dict(launches=cnt, avg_ms=ms / cnt, bytes_per_launch=by / cnt,
                               gbs=(by / cnt) / (ms / cnt * 1e-3) / 1e9)
            total_bytes += by
    st.timing(False)
    # the reference schedule relabels g in place, so g matches the final layout
    ex, n2 = st.expect_cut(g)
    idx, prob = st.topk(8)
    res = dict(workload=name, desc=w["desc"], N=N, P=P, C=C, dtype=dtype, schedule=schedule, steps=steps,
               seconds=secs, layers_per_s=P * steps / secs, ms_per_step=secs / steps * 1e3,
               layer_hbm_frac=(total_bytes / secs) / (HBM_PEAK_GBS * 1e9),
               state_gb=st.nbytes / 1e9, state_bytes=st.nbytes, expect_cut=ex, norm2=n2,
               top8=[[int(i), float(p)] for i, p in zip(idx, prob)], kernels=kern)
    st.destroy()
    g.destroy()
    return res


def cpu_baseline(n30_layers=2):
    """The oracle (the C restatement of the reference's layer loop, complex128,
    OpenMP) on the host cores, in a child process started before this process
    touches the GPU: OMP_PROC_BIND=close, OMP_NUM_THREADS as set (else the
    physical cores this process may use).  Times full layers with a monotonic
    clock around the layer loop alone (src/Weighted.cpp:113,140)."""
    import subprocess

    env = dict(os.environ)
    env["OMP_PROC_BIND"] = "close"
    if not env.get("OMP_NUM_THREADS"):
        env["OMP_NUM_THREADS"] = str(len(os.sched_getaffinity(0)))
    out = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"),
                          "--n30-layers", str(n30_layers)], env=env, capture_output=True, text=True,
                         timeout=900)
    if out.returncode != 0:
        raise RuntimeError(f"cpu_baseline.py failed ({out.returncode}): {out.stderr[-500:]}")
    d = json.loads(out.stdout.strip().splitlines()[-1])
    n30, n20 = d["n30"], d["n20"]
    return dict(value=n30["layers_per_s"], unit="layers/s", cores=d["threads"], kind="port",
                cpu_model=d["cpu_model"], physical_cores_allowed=d["physical_cores_allowed"],
                logical_cpus_allowed=d["logical_cpus_allowed"], omp_proc_bind=d["omp_proc_bind"],
                n20_p5_layers_per_s=n20["layers_per_s"], n20_p5_seconds=n20["seconds"],
                sample=(f"oracle (C restatement of src/Weighted.cpp:97-137, complex128, OpenMP, "
                        f"{d['threads']} threads, OMP_PROC_BIND=close) timed around the layer loop: the first "
                        f"{n30['P']} full layers of N=30 G(30,0.5) C=13 (launch-control layer, then "
                        f"read-modify-write layers) in {n30['seconds']:.1f} s; no scaling. Also the whole "
                        f"N=20 P=5 C=12 loop: {n20['seconds']:.3f} s"))


def serving_bench(n_sat=10000, every=50, n_lat=600, seed=20251207):
    """The serving workload: a stream of 10,000 requests (N uniform in
    {16..30}, P uniform in {1..10}, G(N, p) with p uniform in [0.2, 0.8],
    weights in (0, 1], random angles, 80% re-evaluations of resident graphs),
    complex64, 200 GB HBM budget.

    * saturated: all n_sat requests queued at t=0 through the server;
    * baseline: every `every`-th request of the stream (a systematic sample,
      so its bytes per request match the whole stream's; both are reported)
      one at a time with plain launches (fresh state and graph, the
      reference schedule, <C>, release), and the same sample through the
      server (like for like);
    * latency: n_lat requests arriving as a Poisson stream at 0.8x the
      saturated rate (p50/p99)."""
    from forqaoa.serve import Server, bytes_floor, latency_stats, make_stream, one_at_a_time

    full = make_stream(10000, seed=seed)
    sat = type(full)(graphs=full.graphs, requests=full.requests[:n_sat])
    for r in sat.requests:
        r.arrival = 0.0
    srv = Server(dtype="c64", budget_bytes=200e9)
    t0 = time.perf_counter()
    done = srv.run(sat, realtime=False)
    t_sat = time.perf_counter() - t0
    srv.close()
    evals_s = len(done) / t_sat
    sat_by = {r.rid: r for r in done}
    worst_norm = max(abs(r.norm2 - 1.0) for r in done)

    def copy(reqs):
        return [type(r)(r.rid, r.N, r.P, r.graph_key, r.gammas, r.betas) for r in reqs]

    sample = sat.requests[::every]
    base_reqs = copy(sample)
    t_base = one_at_a_time(full, base_reqs, dtype="c64", schedule="reference")
    worst = max(abs(r.expect - sat_by[r.rid].expect) / abs(sat_by[r.rid].expect) for r in base_reqs)
    sub = type(full)(graphs=full.graphs, requests=copy(sample))
    srv = Server(dtype="c64", budget_bytes=200e9)
    t0 = time.perf_counter()
    srv.run(sub, realtime=False)
    t_sub = time.perf_counter() - t0
    srv.close()

    rate = 0.8 * evals_s
    lat_stream = make_stream(n_lat, seed=seed + 1, rate=rate)
    srv = Server(dtype="c64", budget_bytes=200e9)
    lat_done = srv.run(lat_stream, realtime=True)
    srv.close()
    lat = latency_stats(lat_done)
    floor_all = bytes_floor(full, sat.requests, "c64", "z2")
    floor_sample = bytes_floor(full, sample, "c64", "z2")
    base_rate = len(sample) / t_base
    floor_rate = len(sat.requests) / (floor_all / (HBM_PEAK_GBS * 1e9))
    return dict(
        workload="10,000-request stream (N 16..30, P 1..10, G(N,p) p in [0.2,0.8], 80% resident-graph "
                 f"re-evaluations), complex64, 200 GB budget, half-state slabs; saturated: all of it queued at "
                 f"t=0; baseline: every {every}th request one at a time (reference schedule, fresh state); "
                 f"latency: {n_lat} Poisson arrivals at 0.8x the saturated rate",
        evals_per_s=evals_s, saturated_sample=n_sat, saturated_seconds=t_sat,
        saturated_max_norm2_error=worst_norm,
        bandwidth_floor_evals_per_s=floor_rate, frac_of_bandwidth_floor=evals_s / floor_rate,
        baseline_sample=len(sample), baseline_evals_per_s=base_rate,
        speedup_vs_one_at_a_time=evals_s / base_rate,
        server_evals_per_s_same_sample=len(sample) / t_sub,
        speedup_same_sample=(len(sample) / t_sub) / base_rate,
        bytes_per_request_stream=floor_all / len(sat.requests),
        bytes_per_request_sample=floor_sample / len(sample),
        baseline_max_rel_diff=worst,
        latency_rate_per_s=rate, latency_sample=n_lat, **lat)


def serving_saturated(n, seed=20251207):
    """The first n requests of the serving stream, all queued at t=0, through
    the server (complex64, 200 GB budget, half-state slabs)."""
    from forqaoa.serve import Server, bytes_floor, make_stream

    full = make_stream(10000, seed=seed)
    sub = type(full)(graphs=full.graphs, requests=full.requests[:n])
    for r in sub.requests:
        r.arrival = 0.0
    srv = Server(dtype="c64", budget_bytes=200e9)
    t0 = time.perf_counter()
    done = srv.run(sub, realtime=False)
    secs = time.perf_counter() - t0
    srv.close()
    floor = bytes_floor(full, sub.requests, "c64", "z2")
    return dict(requests=n, seconds=secs, evals_per_s=len(done) / secs,
                bandwidth_floor_evals_per_s=n / (floor / (HBM_PEAK_GBS * 1e9)))


def adjoint_bench(n_sets=16, seed=2026):
    """The adjoint workload: N=26, P=8, seeded 4-regular graph, complex64;
    for each of 16 random angle sets one forward evaluation (half-state
    schedule) and one backward pass (fq_backward, on half states).  Reports gradients/s and
    the reverse pass's achieved bandwidth (4S per pair pass)."""
    import numpy as np
    from forqaoa import FQ_K_ADJOINT, FQ_K_FUSED, Graph, State
    from forqaoa import graphs as G

    w = G.workload("n26_reg4")
    N, P = w["N"], w["P"]
    st = State(N, C=14, dtype="c64")
    g = Graph(N, w["edges"])
    rng = np.random.default_rng(seed)
    sets = [(rng.uniform(0, np.pi, P), rng.uniform(0, np.pi, P)) for _ in range(n_sets)]
    st.run(g, *sets[0], "z2")
    st.backward(g, *sets[0])  # warm-up (allocates the adjoint buffer)
    st.timing(True)
    st.timing_reset()
    t0 = time.perf_counter()
    for gm, bt in sets:
        st.run(g, gm, bt, "z2")
        st.backward(g, gm, bt)
    st.sync()
    secs = time.perf_counter() - t0
    ms_a, cnt_a, by_a = st.timing_get(FQ_K_ADJOINT)
    ms_f, cnt_f, by_f = st.timing_get(FQ_K_FUSED)
    st.timing(False)
    st.destroy()
    g.destroy()
    return dict(workload="n26_reg4: N=26 P=8 4-regular, complex64, 16 angle sets, forward (half-state) + "
                         "backward per set", gradients_per_s=n_sets / secs, ms_per_gradient=secs / n_sets * 1e3,
                adjoint_gbs=by_a / ms_a / 1e6 if ms_a else None, adjoint_ms_per_set=ms_a / n_sets,
                forward_ms_per_set=ms_f / n_sets)


def traffic_from_profiles():
    if not os.path.exists(PROFILE_TRAFFIC):
        return None
    with open(PROFILE_TRAFFIC) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 5 device-to-device copies of a 17.18 GB state (counter calibration)")
    ap.add_argument("--workload", default="n30_gnp05")
    ap.add_argument("--dtype", default="c128")
    ap.add_argument("--schedule", default="fused")
    ap.add_argument("--only", choices=["adjoint", "serving"],
                    help="run one part alone and print its record (for profiling): the adjoint "
                         "workload, or the first --requests requests of the serving stream queued at t=0")
    ap.add_argument("--requests", type=int, default=300)
    args = ap.parse_args()
    if args.only == "adjoint":
        print(json.dumps({"adjoint": adjoint_bench()}), flush=True)
        return
    if args.only == "serving":
        print(json.dumps({"serving_saturated": serving_saturated(args.requests)}), flush=True)
        return
    if args.calibrate_copy:
        from forqaoa import State

        a = State(30, C=13, dtype="c128")
        b = State(30, C=13, dtype="c128")
        a.init_zero()
        for _ in range(5):
            b.clone_from(a)
        b.sync()
        print(json.dumps({"calibrate_copy_bytes": a.nbytes, "copies": 5}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build simulates one state vector on one GPU; multi-GPU state-vector "
              "transposition is not part of it (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    cpu = None
    if not args.no_cpu:
        # before anything here touches the GPU: the oracle runs in a child process
        try:
            cpu = cpu_baseline()
            log(f"cpu baseline: {cpu['value']:.4g} layers/s on {cpu['cores']} threads")
        except Exception as e:
            cpu = dict(error=str(e))
    main_res = measure(args.workload, args.dtype, args.schedule, args.steps, args.warmup)
    log(f"main: {main_res['layers_per_s']:.2f} layers/s, layer HBM frac {main_res['layer_hbm_frac']:.3f}")
    # what simple streaming kernels reach over a buffer of the state's size
    stream = None
    try:
        from forqaoa import stream_probe

        stream = stream_probe(main_res["state_bytes"], 3)
        log(f"stream probe over {stream['bytes'] / 1e9:.2f} GB: in-place {stream['inplace_gbs']:.0f}, "
            f"read {stream['read_gbs']:.0f}, write {stream['write_gbs']:.0f} GB/s")
    except Exception as e:
        stream = dict(error=repr(e))
    extra = []
    if not args.no_extra:
        for name, dt, sch, steps, warm in [("n30_reg3", "c128", "fused", 3, 1),
                                            ("n30_gnp05", "c64", "fused", 3, 1),
                                            ("n30_gnp05", "c128", "reference", 2, 1),
                                            ("n30_gnp05", "c64", "reference", 2, 1),
                                            ("n34_gnp03_int", "c64", "fused", 1, 1),
                                            ("n30_gnp05", "c128", "z2", 3, 1),
                                            ("n30_gnp05", "c64", "z2", 3, 1),
                                            ("n34_gnp03_int", "c64", "z2", 1, 1)]:
            try:
                r = measure(name, dt, sch, steps, warm)
                log(f"extra {name} {dt} {sch}: {r['layers_per_s']:.2f} layers/s")
                r.pop("top8", None)
                extra.append(r)
            except Exception as e:  # report, keep the main line
                extra.append(dict(workload=name, dtype=dt, schedule=sch, error=str(e)))
    adjoint = None
    if not args.no_extra:
        try:
            adjoint = adjoint_bench()
            log(f"adjoint: {adjoint['gradients_per_s']:.1f} gradients/s, pair pass {adjoint['adjoint_gbs']:.0f} GB/s")
        except Exception as e:
            adjoint = dict(error=repr(e))
    serving = None
    if not args.no_serve:
        try:
            serving = serving_bench()
            log(f"serving: {serving['evals_per_s']:.1f} evals/s ({serving['frac_of_bandwidth_floor']:.2f} of the "
                f"bandwidth floor), x{serving['speedup_vs_one_at_a_time']:.2f} vs one-at-a-time, "
                f"p50 {serving['p50_ms']:.1f} ms p99 {serving['p99_ms']:.1f} ms")
        except Exception as e:
            serving = dict(error=repr(e))

    dom_name = "fused" if args.schedule == "fused" else "cost"
    dom = main_res["kernels"].get(dom_name, {})
    traffic = traffic_from_profiles() if args.workload == "n30_gnp05" and args.dtype == "c128" else None
    roof = dict(bound="hbm", achieved=dom.get("gbs"), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=(dom.get("gbs") or 0) / HBM_PEAK_GBS, traffic=traffic,
                kernel=("k_tile (fused tile pass: cost phase + RX through LDS)" if dom_name == "fused"
                        else "k_tile (cost phase)"),
                bytes_per_launch=dom.get("bytes_per_launch"), avg_launch_ms=dom.get("avg_ms"))
    if stream and stream.get("inplace_gbs"):
        # the dominant kernel reads and writes the state: compare with the
        # in-place streaming rate measured on the same device
        roof["stream_inplace_gbs"] = stream["inplace_gbs"]
        roof["frac_of_stream_inplace"] = (dom.get("gbs") or 0) / stream["inplace_gbs"]
    out = {
        "metric": "QAOA layers/s at N=30,34 + %HBM peak; served evals/s with p50/p99 latency",
        "value": main_res["layers_per_s"],
        "unit": "layers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.dtype == "c128" else "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {main_res['desc']}, "
                               f"{args.dtype}, {args.schedule} schedule, "
                               f"one P={main_res['P']} evaluation per step",
                   "N": main_res["N"], "P": main_res["P"], "C": main_res["C"],
                   "schedule": args.schedule, "parallelism": "1 GPU"},
        "layer_hbm_frac": main_res["layer_hbm_frac"],
        "expect_cut": main_res["expect_cut"],
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": main_res["kernels"],
        "extra": extra,
        "adjoint": adjoint,
        "serving": serving,
        "stream_probe": stream,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
