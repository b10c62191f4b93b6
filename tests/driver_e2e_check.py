"""Compare the weighted driver's --dump files and printed <C> (tests/driver_e2e.sh)
with the oracle: states in logical order within 1e-12 (complex128) / 1e-5
(complex64) relative, the .qmap sidecar equal to the reference schedule's
physical->logical map (identity for the fused and half-state schedules), and
the reference's output lines present."""
import math, os, sys
import numpy as np
R = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")
sys.path.insert(0, os.path.join(R, "oracle")); sys.path.insert(0, os.path.join(R, "for-qaoa_amd"))
import pyoracle as O
from forqaoa import graphs as G
out = sys.argv[1]
ok = True
def check(tag, N, C, P, edges, gam, bet, dtype, sched):
    global ok
    sv, qmap, W = O.run_weighted(N, C, edges, gam, bet)
    lg = O.to_logical(sv, qmap)
    npd = np.complex128 if dtype == "c128" else np.complex64
    got = np.fromfile(os.path.join(out, tag + ".bin"), dtype=npd).astype(np.complex128)
    err = np.linalg.norm(got - lg) / np.linalg.norm(lg)
    qm = [int(x) for x in open(os.path.join(out, tag + ".bin.qmap")).read().split()]
    want_qm = list(qmap) if sched == "reference" else list(range(N))
    lines = open(os.path.join(out, tag + ".txt")).read().splitlines()
    ex = float([l for l in lines if l.startswith("Expected cut")][0].split(":")[1])
    ex_ref = O.expect_cut(sv, W)
    tol = 1e-12 if dtype == "c128" else 1e-5
    good = err < tol and qm == want_qm and abs(ex - ex_ref) / ex_ref < tol and lines[0].startswith("Total simulation time")
    ok &= good
    print(f"{tag}: state rel err {err:.2e}, <C> {ex:.12g} vs {ex_ref:.12g}, qmap {'ok' if qm == want_qm else 'MISMATCH'} -> {'ok' if good else 'FAIL'}")
N = 12
check("kn_ref", N, 5, 2, G.complete(N), [math.pi / 4] * 2, [math.pi / 4] * 2, "c128", "reference")
edges = [tuple(map(float, l.split())) for l in open(os.path.join(out, "g.txt")) if len(l.split()) == 3]
edges = [(int(u), int(v), w) for u, v, w in edges]
ang = [tuple(map(float, l.split())) for l in open(os.path.join(out, "a.txt"))]
g_ = [a for a, _ in ang]; b_ = [b for _, b in ang]
for sched in ["reference", "fused", "z2"]:
    for dt in ["c128", "c64"]:
        check(f"file_{sched}_{dt}", 14, 6, 3, edges, g_, b_, dt, sched)
print("ALL OK" if ok else "FAILURES")
sys.exit(0 if ok else 1)
