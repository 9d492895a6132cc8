"""art with xref, the reference's early stop (AIRMs.f90:322, 330-333), on the device against the
compiled reference's fixtures: the stopping iteration, the recorded rows (zero where not reached)
and the iterate before the stop, within 1e-12; the draws consumed by randomized Kaczmarz; and the
Fortran drop-in, whose random_number stream continues right after the draws the reference
consumed."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, assert_close_1e12, load_golden

pytestmark = pytest.mark.gpu


def cases(g):
    return sorted({k.split("/")[0] for k in g.files if k.startswith("case")}, key=lambda c: int(c[4:]))


def options(airms, g, c):
    o = airms.Opts()
    lb, ub = float(g[f"{c}/lb"]), float(g[f"{c}/ub"])
    if np.isfinite(lb):
        o.set_lbound(lb)
    if np.isfinite(ub):
        o.set_ubound(ub)
    o.relaxpar = float(g[f"{c}/relax"])
    o.rowprobability = bool(g[f"{c}/rowprob"])
    return o


def rows_close(X, ref):
    assert X.shape == ref.shape
    for k in range(ref.shape[0]):
        if not ref[k].any():
            assert not X[k].any(), f"row {k} should not be reached"
        else:
            assert_close_1e12(X[k], ref[k])


def test_art_xref_matches_reference_fixtures():
    from fortran_airtools_amd import airms, sparse_matrices as sm
    g = load_golden("art_oracle_small.npz")
    IA, JA, vA, xt = g["IA"], g["JA"], g["vA"], g["xt"]
    n = len(xt)
    nord = int(np.count_nonzero(np.diff(IA)))
    for c in cases(g):
        A = sm.CSRMatrix.from_arrays(len(IA) - 1, n, IA, JA, vA)
        meth = str(g[f"{c}/method"])
        x0 = g[f"{c}/x0"]
        info = {}
        X = airms.art(meth, A, g[f"{c}/b"], g[f"{c}/kin"], x0=x0 if len(x0) else None,
                      options=options(airms, g, c), uniforms=g[f"{c}/u"] if meth == "randkaczmarz" else None,
                      xref=xt, info=info)
        assert info["exit_iter"] == int(g[f"{c}/stop"]), c
        assert info["consumed"] == len(g[f"{c}/u"]), c
        rows_close(X, g[f"{c}/X"])
        if meth == "randkaczmarz":
            assert len(g[f"{c}/u"]) >= 2 * nord


def test_art_xref_needs_the_draws_it_takes():
    from fortran_airtools_amd import airms, sparse_matrices as sm
    from fortran_airtools_amd._lib import FairError
    g = load_golden("art_oracle_small.npz")
    IA, JA, vA, xt = g["IA"], g["JA"], g["vA"], g["xt"]
    A = sm.CSRMatrix.from_arrays(len(IA) - 1, len(xt), IA, JA, vA)
    c = "case4"
    u = g[f"{c}/u"]
    with pytest.raises(FairError, match="too few uniforms"):
        airms.art("randkaczmarz", A, g[f"{c}/b"], g[f"{c}/kin"], options=options(airms, g, c),
                  uniforms=u[:-1], xref=xt)
    # without xref the whole run's draws are required
    with pytest.raises(FairError, match="too few uniforms"):
        airms.art("randkaczmarz", A, g[f"{c}/b"], g[f"{c}/kin"], options=options(airms, g, c), uniforms=u)


def test_art_without_stop_equals_plain_art():
    """An xref run that never stops records exactly what art records."""
    from fortran_airtools_amd import airms, sparse_matrices as sm
    g = load_golden("art_oracle_small.npz")
    IA, JA, vA, xt = g["IA"], g["JA"], g["vA"], g["xt"]
    A = sm.CSRMatrix.from_arrays(len(IA) - 1, len(xt), IA, JA, vA)
    c = "case2"
    assert int(g[f"{c}/stop"]) == 0
    o = options(airms, g, c)
    Xo = airms.art("kaczmarz", A, g[f"{c}/b"], g[f"{c}/kin"], options=o, xref=xt)
    Xp = airms.art("kaczmarz", A, g[f"{c}/b"], g[f"{c}/kin"], options=o)
    assert Xo.tobytes() == Xp.tobytes()


def test_fortran_art_xref_dropin(tmp_path):
    from fortran_airtools_amd import configs as cf
    exe = os.path.join(ROOT, "fortran_airtools_amd", "bin", "art_oracle_check")
    if not os.path.exists(exe):
        pytest.fail("fortran_airtools_amd/bin/art_oracle_check is not built (run __graft_entry__.build())")
    g = load_golden("art_oracle_small.npz")
    geo = cf.Geometry(n=16, p=21, nth=30, deg_step=12.0)
    xt = g["xt"]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([geo.n, geo.p, geo.nth], np.int32).tofile(f)
        np.array([geo.r, geo.dw, geo.sd], np.float64).tofile(f)
        geo.theta().astype(np.float64).tofile(f)
        xt.astype(np.float64).tofile(f)
        g["case3/x0"].astype(np.float64).tofile(f)
        np.array([geo.m], np.int32).tofile(f)
        for c in ("case3", "case4", "case5"):
            g[f"{c}/b"].astype(np.float64).tofile(f)
    r = subprocess.run([exe, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    n, off = len(xt), 0
    for c in ("case3", "case4", "case5"):
        ref = g[f"{c}/X"]
        X = np.frombuffer(raw, np.float64, ref.size, off).reshape(ref.shape)
        off += ref.nbytes
        nxt = int(np.frombuffer(raw, np.int32, 1, off)[0])
        off += 4
        rows_close(X, ref)
        assert nxt == len(g[f"{c}/u"]) + 1, (c, nxt)
    assert off == len(raw)
