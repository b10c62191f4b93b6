"""End-to-end check of the `weighted` driver binary against the oracle.

    python tests/driver_e2e.py OUT_DIR      (needs an MI355X for the runs)

run_driver(out) starts the driver (for-qaoa_amd/host/weighted) as a child
process on
  * the reference's own instance: K_12, unit weights, gamma = beta = pi/4,
    `weighted 2 12 0 5 10` (positionals as src/Weighted.cpp:58-71), reference
    schedule, with --report --dump;
  * a seeded G(14, 0.5) instance read from --graph / --angles files (P=3,
    C=6), with --dump, in every schedule (reference, fused, z2) and both
    amplitude types.
check(out) then compares on the CPU:
  * the --dump state (logical order) with the oracle's final state:
    ||a - b|| / ||b|| < 1e-12 (complex128), 1e-5 (complex64);
  * the .qmap sidecar with the oracle's physical->logical map for the
    reference schedule, the identity for the others;
  * the printed "Expected cut" with the oracle's <C> (same bars);
  * the reference's four output lines (src/Weighted.cpp:153-156) being
    present.

The runs must start from a process that has not touched the GPU (the driver
is a separate GPU process); __graft_entry__.smoke() calls run_driver first.
"""
from __future__ import annotations

import math
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for _p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "for-qaoa_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

WEIGHTED = os.path.join(ROOT, "for-qaoa_amd", "host", "weighted")
FILE_N, FILE_C, FILE_P = 14, 6, 3
KN_N, KN_C, KN_P = 12, 5, 2
REF_LINES = ("Total simulation time", "Total RX time", "Total csqs time", "Total transmission time")


def file_instance():
    from forqaoa import graphs as G

    edges = G.gnp(FILE_N, 0.5, 77, "uniform")
    g, b = G.angles(FILE_P, 78)
    return edges, list(map(float, g)), list(map(float, b))


def runs():
    """(tag, argv after the binary, N, C, edges, gammas, betas, dtype, schedule)"""
    from forqaoa import graphs as G

    out = [("kn_ref", [str(KN_P), str(KN_N), "0", str(KN_C), "10", "--schedule", "reference"], KN_N, KN_C,
            G.complete(KN_N), [math.pi / 4] * KN_P, [math.pi / 4] * KN_P, "c128", "reference")]
    edges, g, b = file_instance()
    for sched in ("reference", "fused", "z2"):
        for dt in ("c128", "c64"):
            out.append((f"file_{sched}_{dt}", [str(FILE_P), str(FILE_N), "0", str(FILE_C), "12", "--graph", "g.txt",
                                               "--angles", "a.txt", "--schedule", sched, "--dtype", dt],
                        FILE_N, FILE_C, edges, g, b, dt, sched))
    return out


def run_driver(out, timeout=60):
    """Writes the instance files and runs the driver once per case, each in a
    child process, with --report --dump <out>/<tag>.bin; stdout to <tag>.txt."""
    os.makedirs(out, exist_ok=True)
    edges, g, b = file_instance()
    with open(os.path.join(out, "g.txt"), "w") as f:
        f.write(f"{FILE_N}\n" + "".join(f"{u} {v} {float(w)!r}\n" for u, v, w in edges))
    with open(os.path.join(out, "a.txt"), "w") as f:
        f.write("".join(f"{x!r} {y!r}\n" for x, y in zip(g, b)))
    for tag, argv, *_ in runs():
        argv = [os.path.join(out, a) if a in ("g.txt", "a.txt") else a for a in argv]
        cmd = [WEIGHTED] + argv + ["--report", "--dump", os.path.join(out, tag + ".bin")]
        print("driver:", " ".join(cmd[1:]), flush=True)
        with open(os.path.join(out, tag + ".txt"), "w") as f:
            rc = subprocess.run(cmd, stdout=f, stderr=subprocess.STDOUT, timeout=timeout).returncode
        if rc != 0:
            with open(os.path.join(out, tag + ".txt")) as f:
                raise RuntimeError(f"weighted {' '.join(argv)} exited with {rc}: {f.read()[-800:]}")


def check_one(out, tag, N, C, edges, gammas, betas, dtype, sched):
    import pyoracle as O

    sv, qmap, W = O.run_weighted(N, C, edges, gammas, betas)
    want = O.to_logical(sv, qmap)
    npd = np.complex128 if dtype == "c128" else np.complex64
    got = np.fromfile(os.path.join(out, tag + ".bin"), dtype=npd).astype(np.complex128)
    err = float(np.linalg.norm(got - want) / np.linalg.norm(want)) if got.size == want.size else math.inf
    with open(os.path.join(out, tag + ".bin.qmap")) as f:
        qm = [int(x) for x in f.read().split()]
    want_qm = [int(q) for q in qmap] if sched == "reference" else list(range(N))
    with open(os.path.join(out, tag + ".txt")) as f:
        lines = f.read().splitlines()
    ex_line = [ln for ln in lines if ln.startswith("Expected cut")]
    ex = float(ex_line[0].split(":")[1]) if ex_line else math.nan
    ex_ref = O.expect_cut(sv, W)
    tol = 1e-12 if dtype == "c128" else 1e-5
    have_lines = all(any(ln.startswith(k) for ln in lines) for k in REF_LINES)
    good = err < tol and qm == want_qm and abs(ex - ex_ref) / abs(ex_ref) < tol and have_lines
    print(f"{tag}: state rel err {err:.2e}, <C> {ex:.12g} vs oracle {ex_ref:.12g}, "
          f"qmap {'ok' if qm == want_qm else 'MISMATCH'}, output lines {'ok' if have_lines else 'MISSING'} "
          f"-> {'ok' if good else 'FAIL'}", flush=True)
    return good


def check(out):
    ok = True
    for tag, _argv, N, C, edges, g, b, dt, sched in runs():
        ok &= check_one(out, tag, N, C, edges, g, b, dt, sched)
    return ok


if __name__ == "__main__":
    d = sys.argv[1]
    run_driver(d)
    good = check(d)
    print("ALL OK" if good else "FAILURES")
    sys.exit(0 if good else 1)
