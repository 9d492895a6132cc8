"""quicksort (system_generation.f90:37-77) in the Fortran drop-in, and the oracle's row steps
against the reference's basic_sweep (AIRMs.f90:813-858), on fixtures from the compiled reference
(tests/golden/make_golden.py --sweep).  Host only."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, bitequal, load_golden


def test_dropin_quicksort_matches_reference(tmp_path):
    exe = os.path.join(ROOT, "fortran_airtools_amd", "bin", "sweep_check")
    if not os.path.exists(exe):
        pytest.fail("fortran_airtools_amd/bin/sweep_check is not built (run __graft_entry__.build())")
    g = load_golden("sweep_small.npz")
    for k in range(5):
        a = g[f"qs{k}/a"]
        fin, fout = tmp_path / f"in{k}.bin", tmp_path / f"out{k}.bin"
        with open(fin, "wb") as f:
            np.array([len(a)], np.int32).tofile(f)
            a.astype(np.float64).tofile(f)
        r = subprocess.run([exe, "sort", str(fin), str(fout)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = fout.read_bytes()
        sa = np.frombuffer(raw, np.float64, len(a), 0)
        si = np.frombuffer(raw, np.int32, len(a), 8 * len(a))
        assert bitequal(sa, g[f"qs{k}/sorted"]), k
        # ties: the same (unstable) order of IND as the reference
        assert bitequal(si, g[f"qs{k}/ind"].astype(np.int32)), k


def sweep_rows(g, name):
    ordv = g[f"{name}/ord"]
    if name == "rowprob":
        return (np.minimum(np.searchsorted(g["cumul"], g[f"{name}/u"], side="left"), len(g["cumul"]) - 1) + 1
                ).astype(np.int32)
    return ordv.astype(np.int32)


@pytest.mark.parametrize("name", ["natural", "reversed", "rowprob"])
def test_oracle_row_steps_match_reference_sweep(oracle, name):
    g = load_golden("sweep_small.npz")
    lb, ub = float(g[f"{name}/lb"]), float(g[f"{name}/ub"])
    x = oracle.art_steps(g["IA"], g["JA"], g["vA"], g["b"], g["nrA"], sweep_rows(g, name), g["x0"],
                         relax=float(g[f"{name}/omg"]), lb=None if np.isnan(lb) else lb,
                         ub=None if np.isnan(ub) else ub)
    assert bitequal(x, g[f"{name}/x"])
