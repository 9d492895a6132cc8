"""ctypes front end of the checkers (test infrastructure only).

`Oracle` wraps liboracle.so, the C restatement in fair_oracle.c.  `Reference` wraps
_ref/libfair_ref.so, the reference's own Fortran compiled by oracle/Makefile, when it has
been built.  Only tests, the smoke check and bench.py's CPU-baseline leg import this module;
the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build_oracle(ref: bool = False) -> None:
    """Compile liboracle.so (and _ref when the reference sources are present)."""
    targets = ["oracle"]
    if ref and os.path.isdir(os.environ.get("REFDIR", "/root/reference")):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def _vp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """C restatement of the reference path (fair_oracle.c)."""

    def __init__(self, path: str | None = None):
        path = path or os.environ.get("FAIR_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build_oracle()
        L = C.CDLL(path)
        self.L = L
        L.oracle_fanlineartomo.restype = C.c_int64
        L.oracle_fanlineartomo.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                           C.c_double, _F64P, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_paralleltomo.restype = C.c_int64
        L.oracle_paralleltomo.argtypes = [C.c_int, C.c_int, C.c_int, _F64P, C.c_void_p,
                                          C.c_void_p, C.c_void_p]
        L.oracle_linspace.argtypes = [C.c_double, C.c_double, C.c_int, _F64P]
        L.oracle_matvec.argtypes = [C.c_int, _I32P, _I32P, _F64P, _F64P, _F64P]
        L.oracle_transvec.argtypes = [C.c_int, C.c_int, _I32P, _I32P, _F64P, _F64P, _F64P]
        L.oracle_transpose.argtypes = [C.c_int, C.c_int, _I32P, _I32P, _F64P, _I32P, _I32P, _F64P]
        L.oracle_row_norms.argtypes = [C.c_int, _I32P, _F64P, C.c_int, _F64P]
        L.oracle_rowvecmult.restype = C.c_double
        L.oracle_rowvecmult.argtypes = [_I32P, _I32P, _F64P, C.c_int, _F64P]
        L.oracle_add_mult_row.argtypes = [_I32P, _I32P, _F64P, C.c_int, C.c_double, _F64P, C.c_int,
                                          C.c_double, C.c_int, C.c_double]
        L.oracle_scalar_mult.argtypes = [C.c_int64, C.c_double, _F64P]
        L.oracle_sirt.argtypes = [C.c_int, C.c_int, C.c_int, _I32P, _I32P, _F64P, _F64P, C.c_int,
                                  C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  _F64P]
        L.oracle_normest.restype = C.c_double
        L.oracle_normest.argtypes = [C.c_int, C.c_int, _I32P, _I32P, _F64P, _F64P, C.c_double,
                                     C.c_int, C.c_void_p]
        L.oracle_gd.restype = C.c_int
        L.oracle_gd.argtypes = [C.c_int, C.c_int, _I32P, _I32P, _F64P, _F64P, C.c_int, C.c_void_p,
                                C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double,
                                C.c_double, _F64P]
        L.oracle_art_steps.argtypes = [_I32P, _I32P, _F64P, _F64P, _F64P, _I32P, C.c_int64,
                                       C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, _F64P]
        L.oracle_scramble.argtypes = [C.c_int, _F64P, _I32P]
        L.oracle_rowprob_cumul.argtypes = [C.c_int, C.c_int, _F64P, _F64P]
        L.oracle_rowprob_pick.restype = C.c_int
        L.oracle_rowprob_pick.argtypes = [C.c_int, _F64P, C.c_double]
        L.oracle_shepplogan.argtypes = [C.c_int, _F64P]
        L.oracle_art.restype = C.c_int64
        L.oracle_art.argtypes = [C.c_int, C.c_int, _I32P, _I32P, _F64P, _F64P, C.c_int, C.c_int,
                                 _I32P, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                 C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _F64P,
                                 C.c_int64, _F64P]
        L.oracle_count_nonempty.restype = C.c_int
        L.oracle_count_nonempty.argtypes = [C.c_int, _I32P]
        L.oracle_twin_art.restype = C.c_int64
        L.oracle_twin_art.argtypes = [C.c_int, C.c_int, _I32P, _I32P, _F64P, _F64P, C.c_int, C.c_int,
                                      C.c_int, _I32P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_int, _F64P, C.c_int64, _F64P,
                                      C.POINTER(C.c_int), _F64P, _F64P, _F64P]

    # -- system_generation -------------------------------------------------------------
    def linspace(self, a, b, n):
        x = np.empty(n)
        self.L.oracle_linspace(a, b, n, x)
        return x

    def fanlineartomo(self, n, p, nth, r, dw, sd, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        m = nth * p
        IA = np.zeros(m + 1, np.int32)
        nnz = self.L.oracle_fanlineartomo(n, p, nth, r, dw, sd, theta, _vp(IA), None, None)
        JA = np.zeros(max(nnz, 1), np.int32)
        vA = np.zeros(max(nnz, 1), np.float64)
        self.L.oracle_fanlineartomo(n, p, nth, r, dw, sd, theta, _vp(IA), _vp(JA), _vp(vA))
        return IA, JA[:nnz], vA[:nnz]

    def paralleltomo(self, n, p, nth, theta):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        m = nth * p
        IA = np.zeros(m + 1, np.int32)
        nnz = self.L.oracle_paralleltomo(n, p, nth, theta, _vp(IA), None, None)
        JA = np.zeros(max(nnz, 1), np.int32)
        vA = np.zeros(max(nnz, 1), np.float64)
        self.L.oracle_paralleltomo(n, p, nth, theta, _vp(IA), _vp(JA), _vp(vA))
        return IA, JA[:nnz], vA[:nnz]

    def shepplogan(self, N):
        x = np.empty(N * N)
        self.L.oracle_shepplogan(N, x)
        return x

    # -- sparse_matrices -----------------------------------------------------------------
    def matvec(self, IA, JA, vA, x):
        y = np.empty(len(IA) - 1)
        self.L.oracle_matvec(len(IA) - 1, IA, _fix32(JA), _fix64(vA), np.ascontiguousarray(x, np.float64), y)
        return y

    def transvec(self, IA, JA, vA, n, y):
        w = np.empty(n)
        self.L.oracle_transvec(len(IA) - 1, n, IA, _fix32(JA), _fix64(vA), np.ascontiguousarray(y, np.float64), w)
        return w

    def transpose(self, IA, JA, vA, n):
        m = len(IA) - 1
        IB = np.empty(n + 1, np.int32)
        JB = np.empty(max(len(JA), 1), np.int32)
        vB = np.empty(max(len(JA), 1), np.float64)
        self.L.oracle_transpose(m, n, IA, _fix32(JA), _fix64(vA), IB, JB, vB)
        return IB, JB[:len(JA)], vB[:len(JA)]

    def row_norms(self, IA, vA, p):
        out = np.empty(len(IA) - 1)
        self.L.oracle_row_norms(len(IA) - 1, IA, _fix64(vA), p, out)
        return out

    def rowvecmult(self, IA, JA, vA, k, v):
        return self.L.oracle_rowvecmult(IA, _fix32(JA), _fix64(vA), k, np.ascontiguousarray(v, np.float64))

    def add_mult_row(self, IA, JA, vA, k, c, x, lb=None, ub=None):
        x = np.array(x, np.float64)
        self.L.oracle_add_mult_row(IA, _fix32(JA), _fix64(vA), k, c, x, lb is not None,
                                   0.0 if lb is None else lb, ub is not None, 0.0 if ub is None else ub)
        return x

    # -- AIRMs --------------------------------------------------------------------------
    def sirt(self, method, IA, JA, vA, n, b, maxits, x0=None, lb=None, ub=None, relax=1.0):
        x0c = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        out = np.empty(n)
        self.L.oracle_sirt({"cimmino": 0, "sart": 1}[method], len(IA) - 1, n, IA, _fix32(JA),
                           _fix64(vA), np.ascontiguousarray(b, np.float64), maxits, _vp(x0c),
                           ub is not None, lb is not None, 0.0 if ub is None else ub,
                           1.0 if lb is None else lb, relax, out)
        return out

    def normest(self, IA, JA, vA, n, v0, tol=1e-5, maxit=100):
        its = C.c_int(0)
        s = self.L.oracle_normest(len(IA) - 1, n, IA, _fix32(JA), _fix64(vA),
                                  np.ascontiguousarray(v0, np.float64), tol, maxit, C.byref(its))
        return s, its.value

    def gd(self, IA, JA, vA, n, b, maxiter, sigma, x0=None, rho=1.5, tol=1e-6, lb=None, ub=None):
        x0c = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        out = np.empty(n)
        its = self.L.oracle_gd(len(IA) - 1, n, IA, _fix32(JA), _fix64(vA),
                               np.ascontiguousarray(b, np.float64), maxiter, _vp(x0c), rho, tol,
                               sigma, ub is not None, lb is not None, 0.0 if ub is None else ub,
                               1.0 if lb is None else lb, out)
        return out, its

    def art_steps(self, IA, JA, vA, b, nrA, rows, x, relax=1.0, lb=None, ub=None):
        x = np.array(x, np.float64)
        rows = np.ascontiguousarray(rows, np.int32)
        self.L.oracle_art_steps(IA, _fix32(JA), _fix64(vA), np.ascontiguousarray(b, np.float64),
                                np.ascontiguousarray(nrA, np.float64), rows, len(rows), relax,
                                ub is not None, lb is not None, 0.0 if ub is None else ub,
                                1.0 if lb is None else lb, x)
        return x

    def art(self, method, IA, JA, vA, n, b, kin, u=None, x0=None, order=None, lb=None, ub=None,
            relax=1.0, damp=0.0, rowprob=False):
        """Returns (Xout[nk, n], uniforms consumed)."""
        kin = np.ascontiguousarray(np.atleast_1d(kin), np.int32)
        out = np.zeros((len(kin), n))
        x0c = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        oc = None if order is None else np.ascontiguousarray(order, np.int32)
        u = np.zeros(1) if u is None else np.ascontiguousarray(u, np.float64)
        q = self.L.oracle_art(len(IA) - 1, n, IA, _fix32(JA), _fix64(vA),
                              np.ascontiguousarray(b, np.float64),
                              {"kaczmarz": 0, "symkaczmarz": 1, "randkaczmarz": 2}[method],
                              len(kin), kin, _vp(x0c), _vp(oc), 0 if oc is None else len(oc),
                              ub is not None, lb is not None, 0.0 if ub is None else ub,
                              1.0 if lb is None else lb, relax, damp, bool(rowprob), u, len(u), out)
        if q == -1:
            raise ValueError("Kin should all be multiples of 2.")
        if q == -2:
            raise ValueError("too few uniforms supplied")
        return out, q

    def twin_art(self, method, stoprule, IA, JA, vA, n, b, kin, u=None, x0=None, y0=None, order=None,
                 slack=7, lb=None, ub=None, relax=1.0, damp=0.0, rowprob=False):
        """twin_art: returns (xout, final_it, Xstore[nk, n], Ystore[nk, n], EG, draws consumed)."""
        kin = np.ascontiguousarray(np.atleast_1d(kin), np.int32)
        nk = len(kin)
        xout = np.zeros(n)
        Xs, Ys = np.zeros((nk, n)), np.zeros((nk, n))
        EG = np.zeros(int(kin.max()))
        fi = C.c_int(0)
        x0c = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        y0c = None if y0 is None else np.ascontiguousarray(y0, np.float64)
        oc = None if order is None else np.ascontiguousarray(order, np.int32)
        u = np.zeros(1) if u is None else np.ascontiguousarray(u, np.float64)
        q = self.L.oracle_twin_art(len(IA) - 1, n, IA, _fix32(JA), _fix64(vA), np.ascontiguousarray(b, np.float64),
                                   {"kaczmarz": 0, "symkaczmarz": 1, "randkaczmarz": 2}[method],
                                   {"errorgauge": 0, "mutualstep": 1}[stoprule], nk, kin, _vp(x0c), _vp(y0c),
                                   _vp(oc), 0 if oc is None else len(oc), slack, ub is not None, lb is not None,
                                   0.0 if ub is None else ub, 1.0 if lb is None else lb, relax, damp, bool(rowprob),
                                   u, len(u), xout, C.byref(fi), Xs, Ys, EG)
        if q == -1:
            raise ValueError("Kin should all be multiples of 2.")
        if q == -2:
            raise ValueError("too few uniforms supplied")
        if q == -3:
            raise ValueError("Initial conditions x0 and y0 too close.")
        return xout, fi.value, Xs, Ys, EG, q

    def count_nonempty(self, IA):
        return self.L.oracle_count_nonempty(len(IA) - 1, IA)

    def scramble(self, nord, u):
        out = np.empty(nord, np.int32)
        self.L.oracle_scramble(nord, np.ascontiguousarray(u, np.float64), out)
        return out

    def rowprob_cumul(self, nrA, nord):
        cum = np.empty(len(nrA))
        self.L.oracle_rowprob_cumul(len(nrA), nord, np.ascontiguousarray(nrA, np.float64), cum)
        return cum

    def rowprob_pick(self, cumul, u):
        return self.L.oracle_rowprob_pick(len(cumul), cumul, u)


def _fix32(a):
    a = np.ascontiguousarray(a, np.int32)
    return a if len(a) else np.zeros(1, np.int32)


def _fix64(a):
    a = np.ascontiguousarray(a, np.float64)
    return a if len(a) else np.zeros(1, np.float64)


class Reference:
    """The reference's own Fortran, compiled into _ref/libfair_ref.so (present only where the
    reference sources were available at build time)."""

    PATH = os.path.join(HERE, "_ref", "libfair_ref.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(cls.PATH)

    def __init__(self):
        L = C.CDLL(self.PATH)
        self.L = L
        i, d = C.c_int, C.c_double
        L.ref_build.restype = i
        L.ref_build.argtypes = [i, i, i, i, i, d, d, d, _F64P]
        L.ref_info.argtypes = [i, C.POINTER(i), C.POINTER(i), C.POINTER(i)]
        L.ref_copy_out.argtypes = [i, i, _I32P, _I32P, _F64P]
        L.ref_set.argtypes = [i, i, i, i, _I32P, _I32P, _F64P]
        L.ref_matvec.argtypes = [i, i, _F64P, i, _F64P]
        L.ref_transvec.argtypes = [i, i, _F64P, i, _F64P]
        L.ref_rowvecmult.argtypes = [i, i, i, _F64P, C.POINTER(d)]
        L.ref_row_norms.argtypes = [i, i, i, _F64P]
        L.ref_transpose.argtypes = [i]
        L.ref_scale.argtypes = [i, d]
        L.ref_add_mult_row.argtypes = [i, i, d, i, _F64P, i, d, i, d]
        L.ref_sirt.argtypes = [i, i, i, _F64P, i, i, i, _F64P, i, i, d, d, d, _F64P]
        L.ref_gd.argtypes = [i, i, _F64P, i, i, i, _F64P, d, d, i, i, d, d, _F64P]
        L.ref_normest.restype = d
        L.ref_normest.argtypes = [i, d, i]
        L.ref_art.argtypes = [i, i, i, _F64P, i, _I32P, i, i, _F64P, i, i, _I32P, i, i, d, d, d, d,
                              i, _F64P]
        L.ref_twin_art.argtypes = [i, i, i, i, _F64P, i, _I32P, i, i, i, _F64P, _F64P, i, i, i, d, d, d, d, i,
                                   _F64P, C.POINTER(i), _F64P, _F64P, _F64P, i]
        L.ref_store.argtypes = [i, C.c_char_p, i]
        L.ref_basic_sweep.argtypes = [i, i, i, _F64P, _F64P, i, _I32P, i, i, d, _F64P, _F64P, i, i, d, d]
        L.ref_quicksort.argtypes = [i, _F64P, _I32P]
        L.ref_art_xref.argtypes = [i, i, i, _F64P, i, _I32P, i, i, _F64P, _F64P, i, i, d, d, d, d, i, _F64P]
        L.ref_rng_seed.argtypes = [i]
        L.ref_set_threads.argtypes = [i]
        L.ref_rng_draw.argtypes = [i, _F64P]
        L.ref_shepplogan.argtypes = [i, _F64P]
        L.ref_phantom.argtypes = [i, i, i, i, d, i, i, _F64P]
        L.ref_white_noise.argtypes = [i, _F64P, d, i, _F64P, _F64P]
        L.ref_norm2_diff.argtypes = [i, _F64P, _F64P, _F64P]
        L.ref_powers.argtypes = [i, _F64P, _F64P, _F64P]
        L.ref_linspace.argtypes = [d, d, i, _F64P]
        L.ref_sum.restype = d
        L.ref_sum.argtypes = [i, _F64P]
        L.ref_read_data.argtypes = [C.c_char_p, i, _I32P, _F64P]
        L.ref_data_arrays.argtypes = [i, _F64P, i, _F64P]
        L.ref_file_mode.argtypes = [i, C.c_char_p, i, i, _I32P, _F64P]
        L.ref_file_mode_b.argtypes = [i, _F64P]
        L.ref_read_ref.argtypes = [C.c_char_p, i, i, _F64P]
        L.ref_read_ordering.argtypes = [C.c_char_p, i, i, C.POINTER(i), _I32P]

    def build(self, slot, kind, n, p, nth, r, dw, sd, theta):
        theta = np.ascontiguousarray(theta, np.float64)
        return self.L.ref_build(slot, kind, n, p, nth, r, dw, sd, theta)

    def info(self, slot):
        m, n, nnz = C.c_int(), C.c_int(), C.c_int()
        self.L.ref_info(slot, C.byref(m), C.byref(n), C.byref(nnz))
        return m.value, n.value, nnz.value

    def get(self, slot):
        m, n, nnz = self.info(slot)
        IA = np.empty(m + 1, np.int32)
        JA = np.empty(max(nnz, 1), np.int32)
        vA = np.empty(max(nnz, 1), np.float64)
        self.L.ref_copy_out(slot, m + 1, IA, JA, vA)
        return IA, JA[:nnz], vA[:nnz]

    def set(self, slot, IA, JA, vA, n):
        m = len(IA) - 1
        self.L.ref_set(slot, m, n, len(JA), np.ascontiguousarray(IA, np.int32), _fix32(JA), _fix64(vA))

    def matvec(self, slot, x):
        m, n, _ = self.info(slot)
        y = np.empty(m)
        self.L.ref_matvec(slot, n, np.ascontiguousarray(x, np.float64), m, y)
        return y

    def transvec(self, slot, y):
        m, n, _ = self.info(slot)
        x = np.empty(n)
        self.L.ref_transvec(slot, m, np.ascontiguousarray(y, np.float64), n, x)
        return x

    def rowvecmult(self, slot, k, v):
        f = C.c_double()
        self.L.ref_rowvecmult(slot, k, len(v), np.ascontiguousarray(v, np.float64), C.byref(f))
        return f.value

    def row_norms(self, slot, p):
        m, _, _ = self.info(slot)
        out = np.empty(m)
        self.L.ref_row_norms(slot, p, m, out)
        return out

    def transpose(self, slot):
        self.L.ref_transpose(slot)

    def scale(self, slot, c):
        self.L.ref_scale(slot, c)

    def add_mult_row(self, slot, k, c, x, lb=None, ub=None):
        x = np.array(x, np.float64)
        self.L.ref_add_mult_row(slot, k, c, len(x), x, lb is not None, 0.0 if lb is None else lb,
                                ub is not None, 0.0 if ub is None else ub)
        return x

    def sirt(self, slot, method, b, maxits, x0=None, lb=None, ub=None, relax=1.0):
        m, n, _ = self.info(slot)
        out = np.empty(n)
        x0a = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64)
        self.L.ref_sirt(slot, {"cimmino": 0, "sart": 1}[method], m, np.ascontiguousarray(b, np.float64),
                        maxits, x0 is not None, n, x0a, ub is not None, lb is not None,
                        0.0 if ub is None else ub, 1.0 if lb is None else lb, relax, out)
        return out

    def gd(self, slot, b, maxiter, x0=None, rho=1.5, tol=1e-6, lb=None, ub=None):
        m, n, _ = self.info(slot)
        out = np.empty(n)
        x0a = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64)
        self.L.ref_gd(slot, m, np.ascontiguousarray(b, np.float64), maxiter, x0 is not None, n, x0a,
                      rho, tol, ub is not None, lb is not None, 0.0 if ub is None else ub,
                      1.0 if lb is None else lb, out)
        return out

    def twin_art(self, slot, method, stoprule, b, kin, x0=None, y0=None, slack=7, lb=None, ub=None, relax=1.0,
                 damp=0.0, rowprob=False):
        """Returns (xout, final_it, Xstore[nk, n], Ystore[nk, n], EG)."""
        m, n, _ = self.info(slot)
        kin = np.ascontiguousarray(np.atleast_1d(kin), np.int32)
        nk = len(kin)
        xout = np.zeros(n)
        Xs, Ys = np.zeros((nk, n)), np.zeros((nk, n))
        neg = int(kin.max())
        EG = np.zeros(neg)
        fi = C.c_int(0)
        x0a = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64)
        y0a = np.zeros(n) if y0 is None else np.ascontiguousarray(y0, np.float64)
        self.L.ref_twin_art(slot, {"kaczmarz": 0, "symkaczmarz": 1, "randkaczmarz": 2}[method],
                            {"errorgauge": 0, "mutualstep": 1}[stoprule], m, np.ascontiguousarray(b, np.float64),
                            nk, kin, x0 is not None, y0 is not None, n, x0a, y0a, slack, ub is not None,
                            lb is not None, 0.0 if ub is None else ub, 1.0 if lb is None else lb, relax, damp,
                            bool(rowprob), xout, C.byref(fi), Xs, Ys, EG, neg)
        return xout, fi.value, Xs, Ys, EG

    def normest(self, slot, tol=1e-5, maxit=100):
        return self.L.ref_normest(slot, tol, maxit)

    def art(self, slot, method, b, kin, x0=None, order=None, lb=None, ub=None, relax=1.0, damp=0.0,
            rowprob=False):
        m, n, _ = self.info(slot)
        kin = np.ascontiguousarray(kin, np.int32)
        out = np.zeros((len(kin), n))
        x0a = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64)
        orda = np.zeros(1, np.int32) if order is None else np.ascontiguousarray(order, np.int32)
        self.L.ref_art(slot, {"kaczmarz": 0, "symkaczmarz": 1, "randkaczmarz": 2}[method], m,
                       np.ascontiguousarray(b, np.float64), len(kin), kin, x0 is not None, n, x0a,
                       order is not None, len(orda), orda, ub is not None, lb is not None,
                       0.0 if ub is None else ub, 1.0 if lb is None else lb, relax, damp,
                       bool(rowprob), out)
        return out  # row k = iterate recorded at kin[k] (Fortran column k)

    def basic_sweep(self, slot, x, b, ord_, nord, nrA, cumul, omg=1.0, isrand=False, rowprob=False, lb=None,
                    ub=None):
        """basic_sweep on a slot's matrix; returns the updated x."""
        m, n, _ = self.info(slot)
        xv = np.array(x, np.float64, copy=True)
        ordv = np.zeros(2 * m, np.int32)
        ordv[:len(ord_)] = ord_
        self.L.ref_basic_sweep(slot, m, n, xv, np.ascontiguousarray(b, np.float64), nord, ordv, bool(isrand),
                               bool(rowprob), omg, np.ascontiguousarray(nrA, np.float64),
                               np.ascontiguousarray(cumul, np.float64), ub is not None, lb is not None,
                               0.0 if ub is None else ub, 0.0 if lb is None else lb)
        return xv

    def quicksort(self, a, ind=None):
        a = np.array(a, np.float64, copy=True)
        ind = np.arange(1, len(a) + 1, dtype=np.int32) if ind is None else np.array(ind, np.int32, copy=True)
        self.L.ref_quicksort(len(a), a, ind)
        return a, ind

    def art_xref(self, slot, method, b, kin, xref, x0=None, lb=None, ub=None, relax=1.0, damp=0.0,
                 rowprob=False):
        """art with the xref early stop; rows of the result not reached stay zero."""
        m, n, _ = self.info(slot)
        kin = np.ascontiguousarray(kin, np.int32)
        out = np.zeros((len(kin), n))
        x0a = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64)
        self.L.ref_art_xref(slot, {"kaczmarz": 0, "symkaczmarz": 1, "randkaczmarz": 2}[method], m,
                            np.ascontiguousarray(b, np.float64), len(kin), kin, x0 is not None, n, x0a,
                            np.ascontiguousarray(xref, np.float64), ub is not None, lb is not None,
                            0.0 if ub is None else ub, 1.0 if lb is None else lb, relax, damp, bool(rowprob), out)
        return out

    def store(self, slot, path):
        p = os.fsencode(path)
        self.L.ref_store(slot, p, len(p))

    def rng_seed(self, seed):
        self.L.ref_rng_seed(seed)

    def set_threads(self, nt):
        """OpenMP threads of the reference's parallel regions from now on."""
        self.L.ref_set_threads(int(nt))

    def rng_draw(self, cnt):
        u = np.empty(cnt)
        self.L.ref_rng_draw(cnt, u)
        return u

    def shepplogan(self, N):
        x = np.empty(N * N)
        self.L.ref_shepplogan(N, x)
        return x

    PHANTOM_KINDS = {"smooth": 1, "threephases": 2, "threephases_smooth": 3, "binary": 4, "fourphases": 5,
                     "mst": 6, "grains": 7, "bubbles": 8}

    def phantom(self, name, N, seed, pin=0, steps=0, v=0.0, gaussian=False):
        """The reference's phantom after rng_seed(seed), vectorised (column-major)."""
        x = np.empty(N * N)
        self.L.ref_phantom(self.PHANTOM_KINDS[name], N, pin, steps, v, int(gaussian), seed, x)
        return x

    def white_noise(self, b, eta, seed):
        """add_white_noise after rng_seed(seed): (bt, e)."""
        b = np.ascontiguousarray(b, np.float64)
        bt, e = np.empty_like(b), np.empty_like(b)
        self.L.ref_white_noise(len(b), b, eta, seed, bt, e)
        return bt, e

    def norm2_diff(self, a, b):
        """NORM2(a(:, k) - b(:, k)) for 2 x cnt arrays given as cnt x 2."""
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        d = np.empty(a.shape[0])
        self.L.ref_norm2_diff(a.shape[0], a, b, d)
        return d

    def linspace(self, a, b, n):
        x = np.empty(n)
        self.L.ref_linspace(a, b, n, x)
        return x

    def fsum(self, v):
        return self.L.ref_sum(len(v), np.ascontiguousarray(v, np.float64))

    @staticmethod
    def _name30(path):
        p = os.fsencode(path)
        if len(p) > 30:
            raise ValueError("the reference's file-name arguments are character(30)")
        return p

    def _data_arrays(self, nth, nt):
        theta, T = np.empty(nth), np.empty(nt)
        self.L.ref_data_arrays(nth, theta, nt, T)
        return theta, T

    def read_data(self, path):
        """read_data: ((n, p, nth), (r, dw, sd), theta, T)."""
        p = self._name30(path)
        ig, rg = np.zeros(3, np.int32), np.zeros(3)
        self.L.ref_read_data(p, len(p), ig, rg)
        n, pp, nth = (int(v) for v in ig)
        return (n, pp, nth), tuple(float(v) for v in rg), *self._data_arrays(nth, nth * pp)

    def file_mode(self, slot, path, skip=0):
        """main.f90's file mode: ((n, p, nth), (r, dw, sd, grid_size), theta, T, bt); A in slot."""
        p = self._name30(path)
        ig, rg = np.zeros(3, np.int32), np.zeros(4)
        self.L.ref_file_mode(slot, p, len(p), skip, ig, rg)
        n, pp, nth = (int(v) for v in ig)
        theta, T = self._data_arrays(nth, nth * pp)
        bt = np.empty(nth * pp)
        self.L.ref_file_mode_b(nth * pp, bt)
        return (n, pp, nth), tuple(float(v) for v in rg), theta, T, bt

    def read_ref(self, path, n):
        p = self._name30(path)
        x = np.empty(n)
        self.L.ref_read_ref(p, len(p), n, x)
        return x

    def read_ordering(self, path, cap):
        p = self._name30(path)
        l = C.c_int()
        out = np.zeros(max(cap, 1), np.int32)
        self.L.ref_read_ordering(p, len(p), cap, C.byref(l), out)
        return out[: min(l.value, cap)], l.value
