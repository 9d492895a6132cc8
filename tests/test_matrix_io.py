"""Matrix dumps in the reference's format (store_matrix / read_csr_matrix,
system_generation.f90:595-707): Fortran unformatted records m, n, nnz, vA, JA, IA.  Host only,
no GPU: the C reader against a file the compiled reference wrote, the C writer against the
flang-built drop-in module's reader and writer, round trips, and the error paths."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, GOLDEN, bitequal, load_golden


def sm():
    from fortran_airtools_amd import sparse_matrices
    return sparse_matrices


def test_reads_the_reference_dump():
    g = load_golden("dump_small.npz")
    m, n, IA, JA, vA = sm().read_matrix_arrays(os.path.join(GOLDEN, "dump_small.bin"))
    assert (m, n) == (int(g["m"]), int(g["n"]))
    assert bitequal(IA, g["IA"]) and bitequal(JA, g["JA"]) and bitequal(vA, g["vA"])


def test_writes_the_reference_bytes(tmp_path):
    """Writing the reference's matrix again gives its file byte for byte."""
    g = load_golden("dump_small.npz")
    out = tmp_path / "again.bin"
    sm().store_matrix_arrays(str(out), int(g["m"]), int(g["n"]), g["IA"], g["JA"], g["vA"])
    with open(os.path.join(GOLDEN, "dump_small.bin"), "rb") as f:
        assert out.read_bytes() == f.read()


def test_roundtrip_random_and_empty(tmp_path):
    rng = np.random.default_rng(3)
    for m, n, nnz in ((7, 5, 20), (3, 9, 0), (1, 1, 1)):
        IA = np.sort(np.concatenate([[0, nnz], rng.integers(0, nnz + 1, m - 1)])).astype(np.int32)
        JA = rng.integers(1, n + 1, nnz).astype(np.int32)
        vA = rng.standard_normal(nnz)
        p = str(tmp_path / f"d{m}.bin")
        sm().store_matrix_arrays(p, m, n, IA, JA, vA)
        mm, nn, IA2, JA2, vA2 = sm().read_matrix_arrays(p)
        assert (mm, nn) == (m, n)
        assert bitequal(IA2, IA) and bitequal(JA2, JA) and bitequal(vA2, vA)


def test_fortran_dropin_reads_and_writes_the_format(tmp_path):
    exe = os.path.join(ROOT, "fortran_airtools_amd", "bin", "matrix_io_check")
    if not os.path.exists(exe):
        pytest.fail("fortran_airtools_amd/bin/matrix_io_check is not built (run __graft_entry__.build())")
    g = load_golden("dump_small.npz")
    mine = tmp_path / "mine.bin"
    sm().store_matrix_arrays(str(mine), int(g["m"]), int(g["n"]), g["IA"], g["JA"], g["vA"])
    back = tmp_path / "back.bin"
    r = subprocess.run([exe, str(mine), str(back)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert back.read_bytes() == mine.read_bytes()


def test_corrupt_and_missing_files(tmp_path):
    from fortran_airtools_amd._lib import FairError
    with pytest.raises(FairError, match="cannot open"):
        sm().read_matrix_arrays(str(tmp_path / "none.bin"))
    with open(os.path.join(GOLDEN, "dump_small.bin"), "rb") as f:
        raw = f.read()
    cut = tmp_path / "cut.bin"
    cut.write_bytes(raw[: len(raw) // 2])
    with pytest.raises(FairError, match="truncated"):
        sm().read_matrix_arrays(str(cut))
    bad = bytearray(raw)
    bad[4 + 4] ^= 0x01                 # trailing marker of the first record
    badp = tmp_path / "bad.bin"
    badp.write_bytes(bytes(bad))
    with pytest.raises(FairError, match="marker"):
        sm().read_matrix_arrays(str(badp))
