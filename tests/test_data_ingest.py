"""Measured-data ingest on the host (no GPU): read_data / read_ref / read_ordering
(system_generation.f90:494-593) and the file mode of main.f90:222-275 (angle cull, grid_size
rescale, A scaled by grid_size/N, bt = -log T).

The fixtures hold what the reference's own read_data, read_ref, read_ordering and file-mode
steps returned for the committed text files (tests/golden/make_golden.py --data).  Where the
compiled reference is present, randomly spelled files are also read by both sides."""
import hashlib
import os
import shutil
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, bitequal, load_golden


def sg():
    from fortran_airtools_amd import system_generation
    return system_generation


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


DATA = os.path.join(GOLDEN, "fan_data_small.txt")


@pytest.fixture(scope="module")
def g():
    return load_golden("fan_data_small.npz")


def test_read_data_matches_reference(g):
    dat = sg().read_data(DATA)
    assert (dat.n, dat.p, dat.nth) == tuple(int(v) for v in g["read/geom"])
    assert bitequal(np.array([dat.r, dat.dw, dat.sd]), g["read/rdwsd"])
    assert bitequal(dat.theta, g["read/theta"])
    assert bitequal(dat.T, g["read/T"])


@pytest.mark.parametrize("skip", [0, 4])
def test_cull_and_rescale_match_reference(g, skip):
    dat = sg().read_data(DATA)
    sg().cull_data(dat, skip)
    grid = sg().rescale_to_grid(dat)
    pre = f"fm{skip}/"
    assert (dat.n, dat.p, dat.nth) == tuple(int(v) for v in g[pre + "geom"])
    assert bitequal(np.array([dat.r, dat.dw, dat.sd, grid]), g[pre + "rg"])
    assert bitequal(dat.theta, g[pre + "theta"])
    assert bitequal(dat.T, g[pre + "T"])
    # the reference's bt = -log(T) is glibc's log: numpy's agrees to the last bit or one ulp
    np.testing.assert_allclose(-np.log(dat.T), g[pre + "bt"], rtol=4e-16, atol=0)


@pytest.mark.parametrize("skip", [0, 4])
def test_oracle_file_mode_matrix_and_sart(oracle, g, skip):
    """Pins the oracle on the file-mode geometry: fanlineartomo at the rescaled geometry times
    grid_size/N gives the reference's matrix bit for bit, and SART on it its iterates."""
    pre = f"fm{skip}/"
    n, p, nth = (int(v) for v in g[pre + "geom"])
    r, dw, sd, grid = (float(v) for v in g[pre + "rg"])
    IA, JA, vA = oracle.fanlineartomo(n, p, nth, r, dw, sd, g[pre + "theta"])
    vA = vA * (grid / n)
    assert bitequal(IA, g[pre + "IA"])
    assert sha(JA) == str(g[pre + "JA_sha"]) and sha(vA) == str(g[pre + "vA_sha"])
    x = oracle.sirt("sart", IA, JA, vA, n * n, g[pre + "bt"], 20, lb=0.0)
    assert bitequal(x, g[pre + "sart_20"])


def test_read_ref_and_ordering(g):
    x = sg().read_ref(os.path.join(GOLDEN, "ref_image_small.txt"), 256)
    assert bitequal(x, g["img/x"])
    order = sg().read_ordering(os.path.join(GOLDEN, "order_small.txt"))
    assert bitequal(order, g["ord/order"].astype(np.int32))


def write(tmp, lines):
    path = os.path.join(tmp, "d.txt")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def header(n=4, p=3, nth=2):
    return [f"n {n}", f"p {p}", f"nth {nth}", "r 2.0", "dw 3.1", "sd 4.0"]


def test_size_checks_raise_the_reference_messages(tmp_path):
    from fortran_airtools_amd._lib import FairError
    bad_angles = header() + ["angles 3", "0", "90", "180", "T 6"] + ["1.0"] * 6
    with pytest.raises(FairError, match="Angle vector has the wrong length, nth is incorrect."):
        sg().read_data(write(str(tmp_path), bad_angles))
    bad_data = header() + ["angles 2", "0", "90", "T 5"] + ["1.0"] * 5
    with pytest.raises(FairError, match="Data has the wrong size, or p and nth are incorrect."):
        sg().read_data(write(str(tmp_path), bad_data))
    short = header() + ["angles 2", "0", "90", "T 6"] + ["1.0"] * 4
    with pytest.raises(FairError, match="end of file"):
        sg().read_data(write(str(tmp_path), short))
    garbage = header() + ["angles 2", "0", "ninety", "T 6"] + ["1.0"] * 6
    with pytest.raises(FairError, match="bad real 'ninety'"):
        sg().read_data(write(str(tmp_path), garbage))
    with pytest.raises(FairError, match="cannot open"):
        sg().read_data(str(tmp_path / "absent.txt"))


def test_load_fan_system_argument_errors(tmp_path):
    from fortran_airtools_amd._lib import FairError
    with pytest.raises(FairError, match="Choose either storing or reading matrix."):
        sg().load_fan_system(DATA, matrixfile=str(tmp_path / "A.bin"), readmatrix=True, storematrix=True)
    with pytest.raises(FairError, match="no matrix file"):
        sg().load_fan_system(DATA, readmatrix=True)


def test_cull_keeps_every_skip_th_angle():
    dat = sg().Data(4, 3, 7, theta=np.arange(7.0), T=np.arange(21.0))
    sg().cull_data(dat, 3)                 # angles 1, 4, 7 (one-based)
    assert dat.nth == 3 and bitequal(dat.theta, np.array([0.0, 3.0, 6.0]))
    assert bitequal(dat.T, np.array([0, 1, 2, 9, 10, 11, 18, 19, 20], np.float64))
    dat = sg().Data(4, 2, 3, theta=np.arange(3.0), T=np.arange(6.0))
    sg().cull_data(dat, 5)                 # skip beyond nth keeps the first angle only
    assert dat.nth == 1 and bitequal(dat.T, np.array([0.0, 1.0]))


# ---- random spellings, read by the compiled reference and by the mirror ---------------------

def spell(rng, v):
    """A list-directed spelling of v that round-trips (17 significant digits)."""
    k = rng.integers(0, 7)
    if k == 0:
        return f"{v:.17e}"
    if k == 1:
        return f"{v:.17g}"
    if k == 2:
        return f"{v:.16E}".replace("E", "D")
    if k == 3:
        return f"{v:.16e}".replace("e", "d")
    if k == 4:
        return f"\t{v:+.17e} ,"
    if k == 5:
        return f"{v:.17e} 12.5 words"
    return f"{v:.16E}".replace("E", "Q") if v != 0 else "0."


def reference():
    from oracle.oracle import Reference
    if not Reference.available():
        pytest.skip("the compiled reference (oracle/_ref) is not built here")
    return Reference()


@pytest.mark.parametrize("seed", range(6))
def test_random_spellings_against_compiled_reference(seed):
    ref = reference()
    rng = np.random.default_rng(100 + seed)
    n, p, nth = int(rng.integers(4, 40)), int(rng.integers(1, 12)), int(rng.integers(1, 9))
    theta = rng.uniform(0, 360, nth)
    T = np.exp(-rng.uniform(0, 4, p * nth))
    labels = ["n", "'N ='", "p,", "\"p\"", "nth:", "x"]
    lines = [f"{labels[rng.integers(0, 2)]} {n}", f"{labels[2 + rng.integers(0, 2)]} {p}",
             f"{labels[4]} {nth}", f"r {spell(rng, 2.0 + rng.random())}", f"dw {spell(rng, 3.0 * rng.random())}",
             f"sd, {spell(rng, 4.0 + rng.random())}", f"angles {nth}"]
    for v in theta:
        lines.append(spell(rng, float(v)))
        if rng.random() < 0.2:
            lines.append("")
    lines.append(f"data {p * nth}")
    lines += [spell(rng, float(v)) for v in T]
    tmp = tempfile.mkdtemp(prefix="fr", dir="/tmp")   # the reference takes character(30) names
    try:
        path = write(tmp, lines)
        geom, rdwsd, th_ref, T_ref = ref.read_data(path)
        dat = sg().read_data(path)
    finally:
        shutil.rmtree(tmp)
    assert (dat.n, dat.p, dat.nth) == geom
    assert bitequal(np.array([dat.r, dat.dw, dat.sd]), np.array(rdwsd))
    assert bitequal(dat.theta, th_ref) and bitequal(dat.T, T_ref)


def test_absorbance_size_arithmetic_without_device():
    """fair_absorbance's size checks run before any device call: no angles with a cull step keeps
    no angles (a null T is then fine), while kept angles with a null T are refused."""
    import ctypes as C
    from fortran_airtools_amd import _lib
    L = _lib.load()
    assert L.fair_absorbance(None, 5, 0, 2, None) == _lib.FAIR_OK
    assert L.fair_absorbance(None, 5, 0, 0, None) == _lib.FAIR_OK
    assert L.fair_absorbance(None, 0, 7, 3, None) == _lib.FAIR_OK
    assert L.fair_absorbance(None, 5, 3, 2, None) == _lib.FAIR_ERR_ARG
    assert L.fair_absorbance_dev(None, 5, 3, 2, None) == _lib.FAIR_ERR_ARG
    assert L.fair_absorbance_dev(None, 5, 0, 2, None) == _lib.FAIR_OK
    assert L.fair_absorbance(None, -1, 3, 2, None) == _lib.FAIR_ERR_ARG
