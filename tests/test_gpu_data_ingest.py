"""The file mode of main.f90:222-275 on the device against the compiled reference's fixtures:
read_data, the angle cull, the grid_size rescale, fanlineartomo scaled by grid_size/N
(IA / JA bit-exact, values within 1e-12), bt = -log T formed on the device (within 1e-15 of
glibc's log), and 20 SART iterations on the result (within 1e-12); the same through the
Fortran drop-in modules."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, assert_close_1e12, bitequal, load_golden

pytestmark = pytest.mark.gpu

DATA = os.path.join(GOLDEN, "fan_data_small.txt")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def check_matrix(IA, JA, vA, g, pre):
    assert bitequal(IA, g[pre + "IA"])
    assert sha(JA) == str(g[pre + "JA_sha"])
    rows, sIA = g[pre + "sample_rows"], g[pre + "sample_IA"]
    sel = np.concatenate([np.arange(IA[k], IA[k + 1]) for k in rows])
    assert bitequal(np.stack([IA[rows], IA[rows + 1]], 1), sIA)
    assert bitequal(JA[sel], g[pre + "sample_JA"])
    assert_close_1e12(vA[sel], g[pre + "sample_vA"])


@pytest.mark.parametrize("skip", [0, 4])
def test_load_fan_system_matches_reference(oracle, skip):
    from fortran_airtools_amd import airms, system_generation as sg
    g = load_golden("fan_data_small.npz")
    pre = f"fm{skip}/"
    A, bt, dat, grid = sg.load_fan_system(DATA, skip=skip)
    assert (dat.n, dat.p, dat.nth) == tuple(int(v) for v in g[pre + "geom"])
    assert bitequal(np.array([dat.r, dat.dw, dat.sd, grid]), g[pre + "rg"])
    IA, JA, vA = A.components()
    check_matrix(IA, JA, vA, g, pre)
    # every value against the oracle (pinned to the reference's matrix hash on the CPU)
    oIA, oJA, ovA = oracle.fanlineartomo(dat.n, dat.p, dat.nth, dat.r, dat.dw, dat.sd, dat.theta)
    assert bitequal(IA, oIA) and bitequal(JA, oJA)
    assert_close_1e12(vA, ovA * (grid / dat.n))
    np.testing.assert_allclose(bt, g[pre + "bt"], rtol=1e-15, atol=0)
    x = airms.sirt("sart", A, bt, [20], options=airms.Opts().set_lbound(0.0))
    assert_close_1e12(x, g[pre + "sart_20"])


def test_absorbance_cull_and_special_values():
    from fortran_airtools_amd import system_generation as sg
    rng = np.random.default_rng(9)
    p, nth = 37, 23
    T = np.exp(-rng.uniform(0, 6, p * nth))
    T[:5] = [1.0, 0.0, -1.0, np.inf, 1e-300]
    for skip in (0, 1, 2, 5, 23, 40):
        b = sg.absorbance(T, p, nth, skip)
        with np.errstate(divide="ignore", invalid="ignore"):
            want = -np.log(T.reshape(nth, p)[:: max(skip, 1)].reshape(-1))
        assert b.shape == want.shape
        assert np.array_equal(np.isnan(b), np.isnan(want)) and np.array_equal(np.isinf(b), np.isinf(want))
        fin = np.isfinite(want)
        np.testing.assert_allclose(b[fin], want[fin], rtol=1e-15, atol=0)
        assert b[0] == 0.0
    assert sg.absorbance(np.zeros(0), 5, 0, 0).shape == (0,)


def test_absorbance_dev_on_hbm():
    import torch
    from fortran_airtools_amd import system_generation as sg
    from fortran_airtools_amd import _lib
    p, nth, skip = 1024, 720, 3
    T = torch.rand(p * nth, dtype=torch.float64, device="cuda") * 0.9 + 0.05
    keep = (nth - 1) // skip + 1
    b = torch.empty(keep * p, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    sg.absorbance_dev(T.data_ptr(), p, nth, skip, b.data_ptr())
    _lib.check(_lib.lib().fair_synchronize())
    want = -torch.log(T.view(nth, p)[::skip].reshape(-1))
    torch.testing.assert_close(b, want, rtol=1e-15, atol=0)


def test_stored_and_read_matrix_file_mode(tmp_path):
    """-s then -r: the stored file-mode matrix read back gives the same system."""
    from fortran_airtools_amd import system_generation as sg
    path = str(tmp_path / "A.bin")
    A, bt, _, _ = sg.load_fan_system(DATA, skip=4, matrixfile=path, storematrix=True)
    B, bt2, _, _ = sg.load_fan_system(DATA, skip=4, matrixfile=path, readmatrix=True)
    assert bitequal(bt, bt2)
    for u, v in zip(A.components(), B.components()):
        assert bitequal(u, v)


def read_stream(path):
    """filemode_check's output (stream access, see the program's header)."""
    raw = open(path, "rb").read()
    off = 0

    def take(dt, cnt):
        nonlocal off
        a = np.frombuffer(raw, dt, cnt, off)
        off += a.nbytes
        return a

    nth, m, nnz = (int(v) for v in take(np.int32, 3))
    rg = take(np.float64, 4)
    out = dict(nth=nth, m=m, nnz=nnz, rg=rg, theta=take(np.float64, nth), bt=take(np.float64, m),
               bd=take(np.float64, m), IA=take(np.int32, m + 1), JA=take(np.int32, nnz), vA=take(np.float64, nnz))
    out["x"] = take(np.float64, 32 * 32)
    if off < len(raw):
        l = int(take(np.int32, 1)[0])
        out["order"] = take(np.int32, l)
        out["xref"] = take(np.float64, 256)
    assert off == len(raw)
    return out


@pytest.mark.parametrize("skip", [0, 4])
def test_fortran_dropin_file_mode(tmp_path, skip):
    exe = os.path.join(ROOT, "fortran_airtools_amd", "bin", "filemode_check")
    if not os.path.exists(exe):
        pytest.fail("fortran_airtools_amd/bin/filemode_check is not built (run __graft_entry__.build())")
    g = load_golden("fan_data_small.npz")
    pre = f"fm{skip}/"
    out = str(tmp_path / "o.bin")
    r = subprocess.run([exe, DATA, str(skip), out, os.path.join(GOLDEN, "order_small.txt"),
                        os.path.join(GOLDEN, "ref_image_small.txt"), "256"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    o = read_stream(out)
    assert o["nth"] == int(g[pre + "geom"][2])
    assert bitequal(o["rg"][[1, 2, 3, 0]], g[pre + "rg"])     # written as grid_size, r, dw, sd
    assert bitequal(o["theta"], g[pre + "theta"])
    np.testing.assert_allclose(o["bt"], g[pre + "bt"], rtol=1e-15, atol=0)
    np.testing.assert_allclose(o["bd"], g[pre + "bt"], rtol=1e-15, atol=0)
    check_matrix(o["IA"], o["JA"], o["vA"], g, pre)
    assert_close_1e12(o["x"], g[pre + "sart_20"])
    assert bitequal(o["order"], g["ord/order"].astype(np.int32))
    assert bitequal(o["xref"], g["img/x"])
