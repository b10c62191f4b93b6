#!/usr/bin/env python3
"""Timed CPU baseline: the oracle's restatement of the reference's layer loop
(src/Weighted.cpp:97-137, complex128, OpenMP) on full layers.

    OMP_PROC_BIND=close OMP_NUM_THREADS=<cores> python oracle/cpu_baseline.py [--n30-layers 2]

TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg runs it as a child
process (so the OpenMP runtime starts with the binding and thread count it
is given, and nothing in it touches a GPU).  The clock brackets the P-layer
loop alone, where MEASURET_START / MEASURET_END sit in the driver
(src/Weighted.cpp:113,140); state allocation and the graph setup are
outside it, as they are in the reference.

Prints one JSON object: layers/s of the first --n30-layers layers of the N=30
G(30, 0.5) instance (C=13, the launch-control layer and full
read-modify-write layers after it), the whole N=20 P=5 loop, the threads
used, and the host's CPU model and physical core count (lscpu).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "for-qaoa_amd"))

import pyoracle as O  # noqa: E402
from forqaoa import graphs as G  # noqa: E402


def host_cpu():
    """CPU model (lscpu's model name, family, model and stepping), physical
    cores and logical CPUs this process may run on.  Called before the
    OpenMP runtime starts: with OMP_PROC_BIND set it pins the calling thread
    to one CPU."""
    model, cores = "unknown", None
    allowed = os.sched_getaffinity(0)
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        info = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                info[k.strip()] = v.strip()
        model = info.get("Model name", "unknown")
        extra = ", ".join(f"{k.lower()} {info[k]}" for k in ("CPU family", "Model", "Stepping") if k in info)
        if extra:
            model = f"{model} ({extra})"
        table = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True,
                               check=True).stdout
        phys = set()
        for line in table.splitlines():
            if line.startswith("#"):
                continue
            cpu, core, sock = (int(x) for x in line.split(","))
            if cpu in allowed:
                phys.add((sock, core))
        cores = len(phys)
    except (OSError, subprocess.CalledProcessError, ValueError):
        pass
    return model, cores, len(allowed)


def timed(name, P):
    w = G.workload(name)
    N, C = w["N"], w["C"]
    _, _, _, secs = O.run_weighted_timed(N, C, w["edges"], w["gammas"][:P], w["betas"][:P])
    return dict(workload=name, N=N, C=C, P=P, seconds=secs, layers_per_s=P / secs)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n30-layers", type=int, default=2)
    args = ap.parse_args()
    model, phys, logical = host_cpu()
    n20 = timed("n20_gnp05", 5)
    n30 = timed("n30_gnp05", args.n30_layers)
    print(json.dumps(dict(n30=n30, n20=n20, threads=O.num_threads(), cpu_model=model,
                          physical_cores_allowed=phys, logical_cpus_allowed=logical,
                          omp_proc_bind=os.environ.get("OMP_PROC_BIND"))), flush=True)


if __name__ == "__main__":
    main()
