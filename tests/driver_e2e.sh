#!/bin/bash
# End-to-end check of the weighted driver (needs an MI355X): runs the driver
# binary on the reference's own instance (K_12, gamma = beta = pi/4, P=2, C=5)
# and on a seeded G(14, 0.5) instance read from --graph/--angles files with every
# schedule and amplitude type, writing the --dump files into OUT; then
# driver_e2e_check.py (CPU) compares them with the oracle.
#
#   bash tests/driver_e2e.sh OUT && python tests/driver_e2e_check.py OUT
#
# The driver is started from this shell, never from a process that already
# uses the GPU.
set -e
O=${1:?output directory}
mkdir -p "$O"
ROOT=$(cd "$(dirname "$0")/.." && pwd)
W=$ROOT/for-qaoa_amd/host/weighted
python3 - "$O" "$ROOT" <<'PY'
import sys
sys.path.insert(0, sys.argv[2] + "/for-qaoa_amd")
from forqaoa import graphs as G
out = sys.argv[1]
e = G.gnp(14, 0.5, 77, "uniform")
g, b = G.angles(3, 78)
open(out + "/g.txt", "w").write("14\n" + "".join(f"{u} {v} {float(w)!r}\n" for u, v, w in e))
open(out + "/a.txt", "w").write("".join(f"{float(x)!r} {float(y)!r}\n" for x, y in zip(g, b)))
PY
timeout -k 5 60 "$W" 2 12 0 5 10 --schedule reference --report --dump "$O/kn_ref.bin" > "$O/kn_ref.txt"
for s in reference fused z2; do
    for d in c128 c64; do
        timeout -k 5 60 "$W" 3 14 0 6 12 --graph "$O/g.txt" --angles "$O/a.txt" --schedule $s --dtype $d \
            --report --dump "$O/file_${s}_${d}.bin" > "$O/file_${s}_${d}.txt"
    done
done
echo "driver runs done"
