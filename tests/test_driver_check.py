"""The driver's end-to-end checker (tests/driver_e2e.py) on the CPU: dumps
written here from the oracle's own results pass it, and each kind of wrong
dump fails it (a perturbed amplitude, a wrong qubit map, a wrong printed <C>,
a missing output line).  The driver runs themselves need a GPU; they happen
in __graft_entry__.smoke()."""
import math

import numpy as np
import pytest

import driver_e2e as D
import pyoracle as O


def _fake_outputs(out, mutate=None):
    for tag, _argv, N, C, edges, g, b, dt, sched in D.runs():
        sv, qmap, W = O.run_weighted(N, C, edges, g, b)
        state = O.to_logical(sv, qmap)
        qm = [int(q) for q in qmap] if sched == "reference" else list(range(N))
        ex = O.expect_cut(sv, W)
        lines = [f"{k} : 0.0" for k in D.REF_LINES] + [f"Expected cut : {ex:.15g}"]
        if mutate and mutate[0] == tag:
            state, qm, lines = mutate[1](state, qm, lines)
        npd = np.complex128 if dt == "c128" else np.complex64
        state.astype(npd).tofile(str(out / f"{tag}.bin"))
        (out / f"{tag}.bin.qmap").write_text(" ".join(map(str, qm)) + "\n")
        (out / f"{tag}.txt").write_text("\n".join(lines) + "\n")


def test_checker_accepts_oracle_outputs(tmp_path):
    _fake_outputs(tmp_path)
    assert D.check(str(tmp_path))


def _bump(state, qm, lines):
    s = state.copy()
    s[5] += 1e-6 * abs(s).max()
    return s, qm, lines


def _swap_qmap(state, qm, lines):
    q = list(qm)
    q[0], q[1] = q[1], q[0]
    return state, q, lines


def _bad_expect(state, qm, lines):
    ex = float(lines[-1].split(":")[1])
    return state, qm, lines[:-1] + [f"Expected cut : {ex * (1 + 1e-9):.15g}"]


def _drop_line(state, qm, lines):
    return state, qm, lines[1:]


@pytest.mark.parametrize("tag,fn", [("file_fused_c128", _bump), ("kn_ref", _swap_qmap),
                                    ("file_z2_c128", _bad_expect), ("file_reference_c64", _drop_line)])
def test_checker_rejects_wrong_outputs(tmp_path, tag, fn):
    _fake_outputs(tmp_path, (tag, fn))
    assert not D.check(str(tmp_path))
