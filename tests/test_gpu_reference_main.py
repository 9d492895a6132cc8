"""The reference's own program (main.f90 with phantomgallery.f90, unmodified) compiled against this
project's Fortran modules and run on the GPU, against the same program built with the reference's
modules, on the same flags.

main.f90 always runs its test problem (main.f90:190-220): paralleltomo with N=256, 400 angles,
384 rays, the phantom of -t (Shepp-Logan by default; here also the other phantoms of the drop-in
phantomgallery module) and add_white_noise at 5e-3 (or -N) from the runtime's random_number,
then the method of -M for -K iterations, writing sol.txt (and errorgauge.txt with -S).  Both
binaries draw from the same runtime generator in the same order, so their outputs agree to the
parity bar (the reference program runs on one thread, see run()).  The binaries are built by
`make -C oracle main` where the reference sources are present; without them the tests skip.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

REF = os.path.join(ROOT, "oracle", "_ref", "main_ref")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "main_dropin")


def run(exe, flags, cwd):
    # one thread: the reference's OpenMP A^T*y scatter (sparse_matrices.f90:382-392) races, so
    # its SIRT results change with the thread count; the device gathers in the one-thread order
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([exe, *flags], cwd=cwd, capture_output=True, text=True, timeout=300, env=env)
    # a Fortran STOP with a message also ends with status 0: require the output file too
    assert r.returncode == 0 and os.path.exists(os.path.join(cwd, "sol.txt")), \
        f"{exe} {flags}: status {r.returncode}; {r.stdout[-2000:]} {r.stderr[-2000:]}"
    out = {"sol": np.loadtxt(os.path.join(cwd, "sol.txt"))}
    eg = os.path.join(cwd, "errorgauge.txt")
    if os.path.exists(eg):
        out["eg"] = np.loadtxt(eg)
    return out


def rel(a, b):
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


CASES = [
    ("sart", ["-Msart", "-K20", "-L0"]),
    ("cimmino", ["-Mcimmino", "-K10", "-L0", "-U1"]),
    ("kaczmarz", ["-Mkaczmarz", "-K2", "-L0", "-U1", "-R0.9"]),
    ("symkaczmarz_errorgauge", ["-Msym", "-K6", "-Seg", "-L0"]),
    ("randkaczmarz", ["-Mrand", "-K2", "-L0"]),
    ("graddescent", ["-Mgd", "-K20", "-L0"]),
    ("kaczmarz_oracle_stop", ["-Mkaczmarz", "-K40", "-Soracle", "-R1.5"]),
    # the other phantoms of -t (the drop-in phantomgallery and add_white_noise on the device)
    ("smooth_sart", ["-tsmooth", "-Msart", "-K5", "-L0"]),
    ("3phases_sart", ["-t3phases", "-Msart", "-K5", "-L0"]),
    ("3phasesmooth_cimmino", ["-t3phasesmooth", "-Mcimmino", "-K5"]),
    ("binary_sart", ["-tbinary", "-Msart", "-K5", "-L0", "-U1"]),
    ("mst_sart", ["-tmst", "-Msart", "-K5"]),
    ("grains_kaczmarz", ["-tgrains", "-Mkaczmarz", "-K1", "-L0"]),
    ("bubbles_sart", ["-tbubbles", "-Msart", "-K5", "-n0.02"]),
]


@pytest.mark.parametrize("name,flags", CASES, ids=[c[0] for c in CASES])
def test_reference_program_on_the_gpu(tmp_path, name, flags):
    if not (os.path.exists(REF) and os.path.exists(DROPIN)):
        pytest.skip("oracle/_ref/main_ref and main_dropin are built only where the reference sources are")
    a, b = tmp_path / "ref", tmp_path / "gpu"
    a.mkdir()
    b.mkdir()
    want = run(REF, flags, str(a))
    got = run(DROPIN, flags, str(b))
    assert got["sol"].shape == want["sol"].shape == (256 * 256,)
    # list-directed output carries 17 significant digits
    assert rel(got["sol"], want["sol"]) <= 1e-12, rel(got["sol"], want["sol"])
    assert set(got) == set(want)
    if "eg" in want and "-Soracle" not in flags:
        # (with -Soracle, art never writes EG: both files hold unset memory)
        # twin_art writes EG only for the iterations it runs (Kin/2 for symkaczmarz); main.f90
        # allocates EG(max_its) without setting it, so the reference's later entries are whatever
        # the memory held, the drop-in's are zero
        k = int(next(f for f in flags if f.startswith("-K"))[2:])
        its = k // 2 if "-Msym" in flags else k
        np.testing.assert_allclose(got["eg"][:its], want["eg"][:its], rtol=1e-12, atol=0)
        assert not got["eg"][its:].any()


def test_reference_program_matrix_dump_and_ordering(tmp_path):
    """main.f90's -s (store_matrix), -r (read_csr_matrix) and -O (read_ordering) paths: both
    programs store their matrix (the dumps must be byte-identical); the drop-in program reads the
    reference's dump and matches the reference program that builds the same matrix (the
    reference's own -r aborts: its read_csr_matrix opens a never-set unit number,
    system_generation.f90:648); Kaczmarz runs on a custom row ordering file.  Solutions compared
    to 1e-12."""
    if not (os.path.exists(REF) and os.path.exists(DROPIN)):
        pytest.skip("oracle/_ref/main_ref and main_dropin are built only where the reference sources are")
    a, b = tmp_path / "ref", tmp_path / "gpu"
    a.mkdir()
    b.mkdir()
    rows = np.random.default_rng(4).choice(400 * 384, 3000, replace=False) + 1
    for d in (a, b):
        (d / "order.txt").write_text("ordering " + str(len(rows)) + "\n" + "\n".join(str(r) for r in rows) + "\n")
    stored = ["-smat.bin", "-Msart", "-K5", "-L0"]
    want, got = run(REF, stored, str(a)), run(DROPIN, stored, str(b))
    assert rel(got["sol"], want["sol"]) <= 1e-12
    da, db = (a / "mat.bin").read_bytes(), (b / "mat.bin").read_bytes()
    assert da == db, "store_matrix dumps differ"
    (b / "mat.bin").write_bytes(da)                 # the reference's dump
    del da, db
    want = run(REF, ["-Mcimmino", "-K5", "-L0"], str(a))
    got = run(DROPIN, ["-rmat.bin", "-Mcimmino", "-K5", "-L0"], str(b))
    assert rel(got["sol"], want["sol"]) <= 1e-12
    ordered = ["-Oorder.txt", "-Mkaczmarz", "-K2", "-L0", "-U1"]
    want, got = run(REF, ordered, str(a)), run(DROPIN, ordered, str(b))
    assert rel(got["sol"], want["sol"]) <= 1e-12
