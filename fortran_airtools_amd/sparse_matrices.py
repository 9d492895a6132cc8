"""Python mirror of the reference module `sparse_matrices` (sparse_matrices.f90).

`CSRMatrix` owns a device CSR (through the C ABI) and exposes the reference's type-bound
procedures under the same names: matvecmul, transvecmul, rowvecmult, row_norms,
column_norms, transpose, add_scalar_mult_row, multiply_by_scalar.  The public components
m, n, nnz, vA, IA, JA read back the device arrays (JA one-based, IA zero-based offsets), as
callers of the reference read them directly (main.f90:288-291).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import FairError, FAIR_ERR_ARG, check, lib, load


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, n=None, what="vector"):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 1 or (n is not None and a.shape[0] != n):
        raise FairError(FAIR_ERR_ARG, "Vector input size incorrect.")
    return a


class CSRMatrix:
    """type(CSR_matrix) (sparse_matrices.f90:157-173) backed by device memory."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        m, n, nnz = C.c_int32(), C.c_int32(), C.c_int64()
        check(lib().fair_csr_info(self._h, C.byref(m), C.byref(n), C.byref(nnz)))
        self._shape = (m.value, n.value, nnz.value)

    # -- construction -----------------------------------------------------------------------
    @classmethod
    def from_arrays(cls, m: int, n: int, IA, JA, vA) -> "CSRMatrix":
        """Upload host components (construct_csr_matrix, sparse_matrices.f90:180-197)."""
        IA = np.ascontiguousarray(IA, np.int32)
        JA = np.ascontiguousarray(JA, np.int32)
        vA = np.ascontiguousarray(vA, np.float64)
        if IA.shape[0] != m + 1 or JA.shape[0] != vA.shape[0]:
            raise FairError(FAIR_ERR_ARG, "inconsistent CSR component sizes")
        h = C.c_void_p()
        check(lib().fair_csr_create(m, n, JA.shape[0], _ptr(IA), _ptr(JA) if len(JA) else None,
                                    _ptr(vA) if len(vA) else None, C.byref(h)))
        return cls(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().fair_csr_free(h)
            except Exception:
                pass
            self._h = None

    def _refresh(self):
        m, n, nnz = C.c_int32(), C.c_int32(), C.c_int64()
        check(lib().fair_csr_info(self._h, C.byref(m), C.byref(n), C.byref(nnz)))
        self._shape = (m.value, n.value, nnz.value)

    # -- public components ------------------------------------------------------------------
    @property
    def m(self) -> int:
        return self._shape[0]

    @property
    def n(self) -> int:
        return self._shape[1]

    @property
    def nnz(self) -> int:
        return self._shape[2]

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def components(self):
        """(IA, JA, vA) as the reference stores them."""
        IA = np.empty(self.m + 1, np.int32)
        JA = np.empty(max(self.nnz, 1), np.int32)
        vA = np.empty(max(self.nnz, 1), np.float64)
        check(lib().fair_csr_download(self._h, _ptr(IA), _ptr(JA), _ptr(vA)))
        return IA, JA[:self.nnz], vA[:self.nnz]

    @property
    def IA(self):
        return self.components()[0]

    @property
    def JA(self):
        return self.components()[1]

    @property
    def vA(self):
        return self.components()[2]

    def device_arrays(self):
        """Device addresses (IA, zero-based JA, vA) as integers."""
        ia, ja, va = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().fair_csr_device_arrays(self._h, C.byref(ia), C.byref(ja), C.byref(va)))
        return ia.value, ja.value, va.value

    # -- type-bound procedures ---------------------------------------------------------------
    def matvecmul(self, v) -> np.ndarray:
        """w = A%matvecmul(v): sparsematmul / MMRS (sparse_matrices.f90:272-298, 349-368)."""
        v = _f64(v)
        w = np.empty(self.m)
        check(lib().fair_matvecmul(self._h, _ptr(v), v.shape[0], _ptr(w), self.m))
        return w

    def transvecmul(self, v) -> np.ndarray:
        """w = A%transvecmul(v): sparsetransmul / MMCS (sparse_matrices.f90:300-347, 370-393)."""
        v = _f64(v)
        w = np.empty(self.n)
        check(lib().fair_transvecmul(self._h, _ptr(v), v.shape[0], _ptr(w), self.n))
        return w

    def rowvecmult(self, k: int, v) -> float:
        """f = A%rowvecmult(k, v): csr_matrowvec (sparse_matrices.f90:252-270), k one-based."""
        v = _f64(v)
        f = C.c_double()
        check(lib().fair_rowvecmult(self._h, int(k), _ptr(v), v.shape[0], C.byref(f)))
        return f.value

    def row_norms(self, p: int) -> np.ndarray:
        """compute_row_norms (sparse_matrices.f90:740-780)."""
        out = np.empty(self.m)
        check(lib().fair_row_norms(self._h, int(p), _ptr(out), self.m))
        return out

    def column_norms(self, p: int) -> np.ndarray:
        """compute_column_norms (sparse_matrices.f90:782-793)."""
        out = np.empty(self.n)
        check(lib().fair_column_norms(self._h, int(p), _ptr(out), self.n))
        return out

    def transpose(self) -> None:
        """call A%transpose(): transposition / reorder (sparse_matrices.f90:395-411, 492-570)."""
        check(lib().fair_transpose(self._h))
        self._refresh()

    def multiply_by_scalar(self, c: float) -> None:
        """scalar_mult (sparse_matrices.f90:824-831)."""
        check(lib().fair_multiply_by_scalar(self._h, float(c)))

    def add_scalar_mult_row(self, k: int, c: float, w, lwrb=None, upb=None) -> np.ndarray:
        """add_mult_row (sparse_matrices.f90:795-822); returns the updated w (the reference
        updates it in place; a float64 ndarray argument is also updated in place here)."""
        inplace = isinstance(w, np.ndarray) and w.dtype == np.float64 and w.flags.c_contiguous
        ww = w if inplace else np.array(w, np.float64)
        if ww.shape[0] != self.n:
            raise FairError(FAIR_ERR_ARG, "Vector input size incorrect.")
        lb = None if lwrb is None else C.byref(C.c_double(lwrb))
        ub = None if upb is None else C.byref(C.c_double(upb))
        check(lib().fair_add_scalar_mult_row(self._h, int(k), float(c), _ptr(ww), ww.shape[0], lb, ub))
        return ww

    # -- device-resident forms ---------------------------------------------------------------
    def matvecmul_dev(self, d_v: int, d_w: int) -> None:
        check(lib().fair_matvecmul_dev(self._h, C.c_void_p(d_v), C.c_void_p(d_w)))

    def transvecmul_dev(self, d_v: int, d_w: int) -> None:
        check(lib().fair_transvecmul_dev(self._h, C.c_void_p(d_v), C.c_void_p(d_w)))


# The reference's CSC tag type has no separate behaviour on the hot path.
sparse_matrix = CSRMatrix
CSR_matrix = CSRMatrix


# ---- matrix dumps (store_matrix / read_csr_matrix, system_generation.f90:595-707) -------------
def store_matrix_arrays(inputfile: str, m: int, n: int, IA, JA, vA) -> None:
    """Write the reference's dump (Fortran unformatted records m, n, nnz, vA, JA, IA); host only."""
    IA = np.ascontiguousarray(IA, np.int32)
    JA = np.ascontiguousarray(JA, np.int32)
    vA = np.ascontiguousarray(vA, np.float64)
    if JA.shape != vA.shape:
        raise FairError(FAIR_ERR_ARG, "JA and vA differ in length")
    L = load()
    check(L.fair_matrix_store(os.fsencode(inputfile), int(m), int(n), int(vA.shape[0]), _ptr(vA), _ptr(JA),
                              _ptr(IA), int(IA.shape[0])))


def read_matrix_arrays(inputfile: str):
    """Read a dump written by store_matrix: returns (m, n, IA, JA, vA); host only."""
    L = load()
    m, n, nia = C.c_int32(), C.c_int32(), C.c_int32()
    nnz = C.c_int64()
    path = os.fsencode(inputfile)
    check(L.fair_matrix_read_header(path, C.byref(m), C.byref(n), C.byref(nnz), C.byref(nia)))
    IA = np.empty(nia.value, np.int32)
    JA = np.empty(max(nnz.value, 1), np.int32)
    vA = np.empty(max(nnz.value, 1), np.float64)
    check(L.fair_matrix_read(path, nnz.value, nia.value, _ptr(vA), _ptr(JA), _ptr(IA)))
    return m.value, n.value, IA, JA[:nnz.value], vA[:nnz.value]


def store_matrix(A: CSRMatrix, inputfile: str) -> None:
    """call store_matrix(A, inputfile), system_generation.f90:595-625."""
    IA, JA, vA = A.components()
    store_matrix_arrays(inputfile, A.m, A.n, IA, JA, vA)


def read_csr_matrix(inputfile: str) -> CSRMatrix:
    """call read_csr_matrix(A, inputfile), system_generation.f90:627-665, into device memory."""
    m, n, IA, JA, vA = read_matrix_arrays(inputfile)
    if IA.shape[0] != m + 1:
        raise FairError(FAIR_ERR_ARG, "the dump is not a CSR matrix (IA has n+1 entries)")
    return CSRMatrix.from_arrays(m, n, IA, JA, vA)

