"""twin_art (AIRMs.f90:371-811) on the device against the compiled reference's fixtures: the
stopping iteration exactly, iterates to 1e-12 (the gauge and the mutual step use device
reductions, whose summation order differs from the reference's sequential NORM2 /
DOT_PRODUCT in the last bits)."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def close(x, ref, tol=1e-12):
    return np.linalg.norm(np.asarray(x) - ref) <= tol * max(np.linalg.norm(ref), 1e-300)


def test_twin_art_matches_reference_fixtures():
    from fortran_airtools_amd import airms, sparse_matrices as sm
    g = load_golden("twin_small.npz")
    IA, JA, vA, b, n = g["IA"], g["JA"], g["vA"], g["b"], len(g["xt"])
    cases = sorted({k.split("/")[0] for k in g.files if k.startswith("case")})
    for c in cases:
        A = sm.CSRMatrix.from_arrays(len(IA) - 1, n, IA, JA, vA)
        lb, ub = float(g[f"{c}/lb"]), float(g[f"{c}/ub"])
        o = airms.Opts()
        if np.isfinite(lb):
            o.set_lbound(lb)
        if np.isfinite(ub):
            o.set_ubound(ub)
        o.relaxpar = float(g[f"{c}/relax"])
        o.damp = float(g[f"{c}/damp"])
        o.rowprobability = bool(g[f"{c}/rowprob"])
        x0, y0 = g[f"{c}/x0"], g[f"{c}/y0"]
        meth = str(g[f"{c}/method"])
        xo, fi, Xs, Ys, EG = airms.twin_art(meth, str(g[f"{c}/stoprule"]), A, b, g[f"{c}/kin"],
                                            x0=x0 if len(x0) else None, y0=y0 if len(y0) else None,
                                            slack=int(g[f"{c}/slack"]), options=o,
                                            uniforms=g[f"{c}/u"] if meth == "randkaczmarz" else None)
        assert fi == int(g[f"{c}/final_it"]), c
        assert close(xo, g[f"{c}/xout"]), c
        assert close(Xs, g[f"{c}/Xstore"]) and close(Ys, g[f"{c}/Ystore"]), c
        assert close(EG, g[f"{c}/EG"]), c


def test_twin_art_rejects_bad_arguments():
    from fortran_airtools_amd import airms, sparse_matrices as sm
    from fortran_airtools_amd._lib import FairError
    g = load_golden("twin_small.npz")
    IA, JA, vA, b, n = g["IA"], g["JA"], g["vA"], g["b"], len(g["xt"])
    A = sm.CSRMatrix.from_arrays(len(IA) - 1, n, IA, JA, vA)
    with pytest.raises(FairError, match="stopping rule"):
        airms.twin_art("kaczmarz", "nostop", A, b, [4])
    with pytest.raises(FairError, match="multiples of 2"):
        airms.twin_art("symkaczmarz", "errorgauge", A, b, [3])
    with pytest.raises(FairError, match="too close"):
        airms.twin_art("kaczmarz", "errorgauge", A, b, [4], x0=np.zeros(n), y0=np.zeros(n))
    with pytest.raises(FairError, match="too few uniforms"):
        airms.twin_art("randkaczmarz", "errorgauge", A, b, [4], uniforms=np.full(10, 0.5))


def test_fortran_twin_art_dropin(tmp_path):
    """The Fortran drop-in's twin_art (fortran/examples/twin_check.f90) draws from the runtime's
    random_number as the reference does: results equal the fixtures, and the generator continues
    right after the draws the reference consumed."""
    import os
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "fortran_airtools_amd", "bin", "twin_check")
    if not os.path.exists(exe):
        pytest.fail("fortran_airtools_amd/bin/twin_check is not built (run __graft_entry__.build())")
    g = load_golden("twin_small.npz")
    n_img, p, nth = 16, 23, 30
    theta = (0.37 + 12.0 * np.arange(nth)) / 180.0 * np.pi
    b = g["b"]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([n_img, p, nth], np.int32).tofile(f)
        np.array([2.0, 3.1, 4.0], np.float64).tofile(f)
        theta.astype(np.float64).tofile(f)
        np.array([len(b)], np.int32).tofile(f)
        b.astype(np.float64).tofile(f)
    r = subprocess.run([exe, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    n = n_img * n_img
    rec = 4 + 8 * n + 4
    assert len(raw) == 3 * rec
    for k, case in enumerate(("case1", "case4", "case5")):
        chunk = raw[k * rec:(k + 1) * rec]
        fi = int(np.frombuffer(chunk, np.int32, 1, 0)[0])
        xo = np.frombuffer(chunk, np.float64, n, 4)
        nxt = int(np.frombuffer(chunk, np.int32, 1, 4 + 8 * n)[0])
        assert fi == int(g[f"{case}/final_it"]), case
        assert close(xo, g[f"{case}/xout"]), case
        used = len(g[f"{case}/u"]) if str(g[f"{case}/method"]) == "randkaczmarz" else 0
        assert nxt == used + 1, (case, nxt, used)
