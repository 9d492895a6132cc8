"""The Fortran drop-in modules (sparse_matrices, system_generation, AIRMs) driven by a
reference-style Fortran program (fortran/examples/dropin_check.f90) on the N=64 problem."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, assert_close_1e12, bitequal, load_golden

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "fortran_airtools_amd", "bin", "dropin_check")


def write_input(path, gold):
    from fortran_airtools_amd import configs as cf
    g = cf.C1
    with open(path, "wb") as f:
        np.array([g.n, g.p, g.nth], np.int32).tofile(f)
        np.array([g.r, g.dw, g.sd], np.float64).tofile(f)
        g.theta().astype(np.float64).tofile(f)
        np.array([g.m], np.int32).tofile(f)
        gold["b"].astype(np.float64).tofile(f)
        gold["probe_x"].astype(np.float64).tofile(f)


def test_fortran_dropin_stops_on_allocation_failure(tmp_path):
    """FAIR_FAIL_ALLOC=1 fails the program's first device allocation (the matrix build): the
    drop-in stops with the library's message, as the reference's `stop "<message>"` does, and
    writes no results."""
    if not os.path.exists(EXE):
        pytest.fail("fortran_airtools_amd/bin/dropin_check is not built (run __graft_entry__.build())")
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, load_golden("c1.npz"))
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, FAIR_FAIL_ALLOC="1"))
    out = r.stdout + r.stderr
    assert "out of memory (allocation 1 failed by FAIR_FAIL_ALLOC)" in out, out
    assert not fout.exists() or fout.stat().st_size == 0


def test_fortran_dropin_c1(tmp_path, oracle):
    if not os.path.exists(EXE):
        pytest.fail("fortran_airtools_amd/bin/dropin_check is not built (run __graft_entry__.build())")
    gold = load_golden("c1.npz")
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, gold)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    m, n, nnz = take(np.int32, 3)
    IA, JA, vA = take(np.int32, m + 1), take(np.int32, nnz), take(np.float64, nnz)
    y, w, X, xc, xs, JA2 = (take(np.float64, m), take(np.float64, n), take(np.float64, n),
                            take(np.float64, n), take(np.float64, n), take(np.int32, nnz))
    assert off == len(raw)
    import hashlib
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    assert int(nnz) == int(gold["nnz"])
    assert bitequal(IA, gold["IA"]) and sha(JA) == str(gold["JA_sha"]) and sha(vA) == str(gold["vA_sha"])
    assert_close_1e12(y, gold["probe_Ax"])
    assert_close_1e12(w, oracle.transvec(IA, JA, vA, n, gold["b"]))
    Xo, _ = oracle.art("symkaczmarz", IA, JA, vA, n, gold["b"], [4], lb=0.0, ub=1.0)
    assert_close_1e12(X, Xo[0])
    assert_close_1e12(xc, gold["cimmino_100"])
    assert_close_1e12(xs, gold["sart_100"])
    # SART leaves the caller's rows sorted by column, as the reference's double transpose does
    IB, JB, vB = oracle.transpose(IA, JA, vA, n)
    _, JS, _ = oracle.transpose(IB, JB, vB, m)
    assert bitequal(JA2, JS)
