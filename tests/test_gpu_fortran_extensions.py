"""fair_extensions driven from Fortran (fortran/examples/extensions_check.f90): the Fortran
wrappers hand the same arrays to the library as the Python mirror does, so FP32 sirt, batched
sirt and the projector's products agree with the Python calls bit for bit; gradient descent
(sigma from a random_number start) agrees to the normest tolerance."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, relerr

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "fortran_airtools_amd", "bin", "extensions_check")


def test_fortran_extensions_c1(tmp_path):
    from fortran_airtools_amd import airms, configs as cf, system_generation as sg
    from fortran_airtools_amd.matrix_free import FanProjector
    if not os.path.exists(EXE):
        pytest.fail("fortran_airtools_amd/bin/extensions_check is not built (run __graft_entry__.build())")
    g = cf.C1
    dat = sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta())
    A = sg.fanlineartomo(dat)
    k = 3
    X = np.stack([cf.shepplogan(g.n), cf.bubbles(g.n), cf.smooth(g.n)], 1)
    B = np.stack([cf.add_noise(A.matvecmul(X[:, j]), 0.01, 20 + j) for j in range(k)], 1).astype(np.float32)
    probe = np.random.default_rng(3).random(g.n ** 2).astype(np.float32)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([g.n, g.p, g.nth], np.int32).tofile(f)
        np.array([g.r, g.dw, g.sd], np.float64).tofile(f)
        g.theta().astype(np.float64).tofile(f)
        np.array([g.m, k], np.int32).tofile(f)
        B.T.astype(np.float32).tofile(f)            # Fortran B(m, k): column-major
        probe.tofile(f)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    n, m = g.n ** 2, g.m
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    xs = take(np.float32, n)
    Xf = take(np.float32, n * k).reshape(k, n).T
    y, w = take(np.float32, m), take(np.float32, n)
    nseg = int(take(np.int64, 1)[0])
    xg = take(np.float32, n)
    assert off == len(raw)
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    assert xs.tobytes() == airms.sirt_f32("sart", A, B[:, 0], [20], options=o).tobytes()
    assert Xf.tobytes() == np.ascontiguousarray(airms.sirt_batch("sart", A, B, [20], options=o)).tobytes()
    P = FanProjector(dat)
    assert y.tobytes() == P.matvecmul(probe).tobytes()
    assert w.tobytes() == P.transvecmul(B[:, 0]).tobytes()
    assert nseg == P.segments()
    import torch
    db = torch.from_numpy(B[:, 0].copy()).cuda()
    dx = torch.zeros(n, dtype=torch.float32, device="cuda")
    P.gradient_descent_dev(db.data_ptr(), dx.data_ptr(), max_iter=10, tolerance=0.0, options=o)
    assert relerr(xg, dx.cpu().numpy()) <= 1e-4
