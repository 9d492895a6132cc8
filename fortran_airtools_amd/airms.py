"""Python mirror of the reference module `AIRMs` (AIRMs.f90).

`Opts` is `type(opts)` (:6-57) with initialise / set_ubound / set_lbound; `sirt`,
`gradient_descent`, `normest` and `art` keep the reference's argument meaning.  Where the
reference draws from the compiler's RANDOM_NUMBER (normest's start vector, randomized
Kaczmarz), the caller supplies the draws: `v0` for normest, `uniforms` for art, in the order
the reference consumes them (`art_uniform_count` tells how many).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FairError, FAIR_ERR_ARG, FAIR_ERR_METHOD, check, lib
from .sparse_matrices import CSRMatrix, _f64, _ptr

SIRT_METHODS = {"cimmino": 0, "sart": 1, "sirt": 1}
ART_METHODS = {"kaczmarz": 0, "symkaczmarz": 1, "sym": 1, "randkaczmarz": 2, "rand": 2}


class Opts:
    """type(opts), AIRMs.f90:6-57."""

    def __init__(self):
        self.initialise()

    def initialise(self):
        """initialise_options (AIRMs.f90:17-37): note the placeholder ubound=0, lbound=1."""
        self.upb = False
        self.lwrb = False
        self.ubound = 0.0
        self.lbound = 1.0
        self.damp = 0.0
        self.relaxpar = 1.0
        self.verbose = False
        self.rowprobability = False
        return self

    def set_ubound(self, ub: float):
        self.upb = True
        self.ubound = float(ub)
        return self

    def set_lbound(self, lb: float):
        self.lwrb = True
        self.lbound = float(lb)
        return self

    def _c(self) -> _lib.Opts:
        return _lib.Opts(float(self.relaxpar), float(self.ubound), float(self.lbound), float(self.damp),
                         int(bool(self.verbose)), int(bool(self.upb)), int(bool(self.lwrb)),
                         int(bool(self.rowprobability)))


def _opts(options):
    return (options or Opts())._c()


def _maxval(Kin) -> int:
    return int(np.max(np.atleast_1d(Kin)))


def sirt(sirt_method: str, A: CSRMatrix, b, Kin, x0=None, options: Opts | None = None) -> np.ndarray:
    """call sirt(xout, sirt_method, A, b, Kin, x0, options), AIRMs.f90:997-1127.  Runs
    maxval(Kin) iterations; SART leaves A's rows sorted by column, as the reference does."""
    if sirt_method not in SIRT_METHODS:
        raise FairError(FAIR_ERR_METHOD, "Not a recognised SIRT method.")
    b = _f64(b, A.m)
    x0a = None if x0 is None else _f64(x0, A.n)
    if x0 is not None and x0a.shape[0] != A.n:
        raise FairError(FAIR_ERR_ARG, "Size of x0 should be equal to n (#columns).")
    out = np.empty(A.n)
    o = _opts(options)
    check(lib().fair_sirt(A.handle, SIRT_METHODS[sirt_method], _ptr(b), _maxval(Kin), _ptr(x0a),
                          C.byref(o), _ptr(out)))
    return out


def sirt_dev(sirt_method: str, A: CSRMatrix, d_b: int, maxits: int, d_x: int,
             options: Opts | None = None) -> None:
    """Device-resident SIRT on HBM vectors (d_x: start in, result out)."""
    o = _opts(options)
    check(lib().fair_sirt_dev(A.handle, SIRT_METHODS[sirt_method], C.c_void_p(d_b), int(maxits),
                              C.c_void_p(d_x), C.byref(o)))


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None and a.shape[0] != n:
        raise FairError(FAIR_ERR_ARG, "Vector input size incorrect.")
    return a


def sirt_f32(sirt_method: str, A: CSRMatrix, b, Kin, x0=None, options: Opts | None = None) -> np.ndarray:
    """Precision variant of sirt: FP32 matrix values, iterate and accumulation, FP64 weights."""
    if sirt_method not in SIRT_METHODS:
        raise FairError(FAIR_ERR_METHOD, "Not a recognised SIRT method.")
    b = _f32(b, A.m)
    x0a = None if x0 is None else _f32(x0, A.n)
    out = np.empty(A.n, np.float32)
    o = _opts(options)
    check(lib().fair_sirt_f32(A.handle, SIRT_METHODS[sirt_method], _ptr(b), _maxval(Kin), _ptr(x0a),
                              C.byref(o), _ptr(out)))
    return out


def sirt_batch(sirt_method: str, A: CSRMatrix, B, Kin, X0=None, options: Opts | None = None,
               use_graph: bool = False) -> np.ndarray:
    """Batched FP32 sirt over the columns of B (shape (m, k)); returns X of shape (n, k)."""
    if sirt_method not in SIRT_METHODS:
        raise FairError(FAIR_ERR_METHOD, "Not a recognised SIRT method.")
    B = np.ascontiguousarray(B, dtype=np.float32)
    if B.ndim != 2 or B.shape[0] != A.m:
        raise FairError(FAIR_ERR_ARG, "B must have shape (m, k)")
    k = B.shape[1]
    X0a = None if X0 is None else np.ascontiguousarray(X0, dtype=np.float32)
    if X0a is not None and X0a.shape != (A.n, k):
        raise FairError(FAIR_ERR_ARG, "X0 must have shape (n, k)")
    out = np.empty((A.n, k), np.float32)
    o = _opts(options)
    check(lib().fair_sirt_batch(A.handle, SIRT_METHODS[sirt_method], _ptr(B), k, _maxval(Kin), _ptr(X0a),
                                C.byref(o), int(bool(use_graph)), _ptr(out)))
    return out


def sirt_batch_dev(sirt_method: str, A: CSRMatrix, d_B: int, k: int, maxits: int, d_X: int,
                   options: Opts | None = None, use_graph: bool = False) -> None:
    """Batched FP32 sirt on HBM panels: d_B (m x k) and d_X (n x k, start in, result out), both
    row-major FP32.  With use_graph the run is captured once and replayed on later calls with
    the same arguments."""
    if sirt_method not in SIRT_METHODS:
        raise FairError(FAIR_ERR_METHOD, "Not a recognised SIRT method.")
    o = _opts(options)
    check(lib().fair_sirt_batch_dev(A.handle, SIRT_METHODS[sirt_method], C.c_void_p(d_B), int(k), int(maxits),
                                    C.c_void_p(d_X), C.byref(o), int(bool(use_graph))))


def autotune_batch(A: CSRMatrix, k: int) -> tuple:
    """Times the batched SpMM's row orders with k right-hand sides and keeps the fastest for A*X
    and A^T*R; returns the two order codes."""
    out = (C.c_int32 * 2)()
    check(lib().fair_autotune_batch(A.handle, int(k), out))
    return out[0], out[1]


def batch_relerr_dev(d_X: int, d_Xref: int, n: int, k: int) -> np.ndarray:
    """Per column ||X_j - Xref_j|| / ||Xref_j|| of an FP32 panel X against an FP64 panel Xref
    (both n x k row-major on the device), with FP64 sums."""
    err = np.empty(int(k))
    check(lib().fair_batch_relerr_dev(C.c_void_p(d_X), C.c_void_p(d_Xref), int(n), int(k), _ptr(err)))
    return err


def normest(A: CSRMatrix, tolerance: float = 1e-5, max_its: int = 100, v0=None, seed: int = 0) -> float:
    """normest (AIRMs.f90:1129-1173): power iteration on A^T A from v0 (uniform [0,1) draws,
    seeded here when not supplied)."""
    v0 = np.random.default_rng(seed).random(A.n) if v0 is None else _f64(v0, A.n)
    s = C.c_double()
    check(lib().fair_normest(A.handle, _ptr(v0), float(tolerance), int(max_its), C.byref(s)))
    return s.value


def gradient_descent(A: CSRMatrix, b, max_iter: int = 100, x0=None, rho_in: float = 1.5,
                     tolerance: float = 1e-6, options: Opts | None = None, sigma: float | None = None,
                     v0=None, return_iters: bool = False):
    """gradient_descent (AIRMs.f90:860-995).  sigma defaults to normest(A, v0=v0)."""
    if tolerance < 0:
        raise FairError(FAIR_ERR_ARG, "Must use positive tolerance.")
    b = _f64(b, A.m)
    x0a = None if x0 is None else _f64(x0, A.n)
    if sigma is None:
        sigma = normest(A, v0=v0)
    out = np.empty(A.n)
    its = C.c_int32()
    o = _opts(options)
    check(lib().fair_gradient_descent(A.handle, _ptr(b), int(max_iter), _ptr(x0a), float(rho_in),
                                      float(tolerance), float(sigma), C.byref(o), _ptr(out), C.byref(its)))
    return (out, its.value) if return_iters else out


def _art_args(art_method, Kin, order):
    if art_method not in ART_METHODS:
        raise FairError(FAIR_ERR_METHOD, "Unknown ART method. Use kaczmarz, symkaczmarz or randkaczmarz")
    K = np.ascontiguousarray(np.atleast_1d(Kin), np.int32)
    ordv = None if order is None else np.ascontiguousarray(order, np.int32)
    return ART_METHODS[art_method], K, ordv


def art_uniform_count(art_method: str, A: CSRMatrix, Kin, order=None, options: Opts | None = None) -> int:
    meth, K, ordv = _art_args(art_method, Kin, order)
    o = _opts(options)
    cnt = C.c_int64()
    check(lib().fair_art_uniform_count(A.handle, meth, K.shape[0], _ptr(K), _ptr(ordv),
                                       0 if ordv is None else ordv.shape[0], C.byref(o), C.byref(cnt)))
    return cnt.value


def art(art_method: str, A: CSRMatrix, b, Kin, x0=None, order=None, options: Opts | None = None,
        uniforms=None, seed: int = 0, xref=None, info: dict | None = None) -> np.ndarray:
    """call art(Xout, art_method, A, b, Kin, x0, xref, order, options), AIRMs.f90:59-369.
    Returns Xout as an array of shape (size(Kin), n): row k is the iterate recorded at Kin(k).
    With xref the run stops once ||x - xref|| grows over an iteration and the last row holds the
    iterate before that iteration (:322, 330-333); rows not reached are zero.  `uniforms` are the
    RANDOM_NUMBER draws of randomized Kaczmarz in the reference's order; seeded uniform draws are
    used when not supplied.  `info`, when given, receives 'exit_iter' (0: no early stop) and
    'consumed' (draws used) of an xref run."""
    meth, K, ordv = _art_args(art_method, Kin, order)
    b = _f64(b, A.m)
    x0a = None if x0 is None else _f64(x0, A.n)
    o = _opts(options)
    if uniforms is None and meth == 2:
        uniforms = np.random.default_rng(seed).random(art_uniform_count(art_method, A, Kin, order, options))
    u = None if uniforms is None else np.ascontiguousarray(uniforms, np.float64)
    out = np.zeros((K.shape[0], A.n))
    nord = 0 if ordv is None else ordv.shape[0]
    if xref is None:
        check(lib().fair_art(A.handle, meth, _ptr(b), K.shape[0], _ptr(K), _ptr(x0a), _ptr(ordv), nord,
                             C.byref(o), _ptr(u), 0 if u is None else u.shape[0], _ptr(out)))
        return out
    xr = _f64(xref, A.n)
    ex, used = C.c_int32(), C.c_int64()
    check(lib().fair_art_oracle(A.handle, meth, _ptr(b), K.shape[0], _ptr(K), _ptr(x0a), _ptr(xr), _ptr(ordv),
                                nord, C.byref(o), _ptr(u), 0 if u is None else u.shape[0], _ptr(out),
                                C.byref(ex), C.byref(used)))
    if info is not None:
        info["exit_iter"], info["consumed"] = ex.value, used.value
    return out


def basic_sweep(x, A: CSRMatrix, b, ord_, nord: int, isrand: bool, rowprob: bool, omg: float, nrA, cumul,
                upb: bool, lwrb: bool, ub: float, lb: float, uniforms=None, seed: int = 0) -> np.ndarray:
    """call basic_sweep(x, A, b, ord, nord, isrand, rowprob, omg, nrA, cumul, upb, lwrb, ub, lb),
    AIRMs.f90:813-858.  Returns the updated x.  With isrand and rowprob the row of each step is the
    first whose cumul reaches that step's RANDOM_NUMBER draw (`uniforms`, nord of them in the
    reference's order; seeded draws when not supplied); otherwise the rows are ord(1:nord)."""
    xs = np.array(_f64(x, A.n), copy=True)
    b = _f64(b, A.m)
    nrA = _f64(nrA, A.m)
    if isrand and rowprob:
        cumul = _f64(cumul, A.m)
        u = np.random.default_rng(seed).random(nord) if uniforms is None else np.asarray(uniforms, np.float64)
        if u.shape[0] < nord:
            raise FairError(FAIR_ERR_ARG, "too few uniforms for the sweep")
        rows = np.minimum(np.searchsorted(cumul, u[:nord], side="left") + 1, A.m).astype(np.int32)
    else:
        rows = np.ascontiguousarray(np.asarray(ord_, np.int32)[:nord])
        if rows.shape[0] < nord:
            raise FairError(FAIR_ERR_ARG, "ord holds fewer than nord rows")
    o = Opts()
    o.relaxpar = float(omg)
    if upb:
        o.set_ubound(ub)
    if lwrb:
        o.set_lbound(lb)
    co = _opts(o)
    check(lib().fair_art_steps(A.handle, _ptr(rows), int(nord), _ptr(b), _ptr(nrA), C.byref(co), _ptr(xs)))
    return xs


def level_schedule(m: int, n: int, IA, JA, seq):
    """The level schedule the Kaczmarz sweeps run a row sequence in (host work, no device call):
    returns (rows, off), the sequence's one-based rows sorted by (level, position) and the level
    offsets.  Steps of one level touch disjoint columns; every column sees its steps in sequence
    order, so the sweep's result does not change.  IA zero-based offsets, JA one-based columns."""
    IA = np.ascontiguousarray(IA, np.int32)
    JA = np.ascontiguousarray(JA, np.int32)
    seq = np.ascontiguousarray(seq, np.int32)
    if IA.shape[0] != m + 1:
        raise FairError(FAIR_ERR_ARG, "IA must hold m + 1 offsets")
    if JA.shape[0] < IA[m]:
        raise FairError(FAIR_ERR_ARG, "JA holds fewer than IA(m) entries")
    rows = np.zeros(max(seq.shape[0], 1), np.int32)
    off = np.zeros(seq.shape[0] + 1, np.int32)
    nlev = C.c_int32()
    # host work only: the library is loaded but no device is initialised
    check(_lib.load().fair_level_schedule(m, n, _ptr(IA), _ptr(JA), _ptr(seq), seq.shape[0], _ptr(rows),
                                          _ptr(off), C.byref(nlev)))
    return rows[: seq.shape[0]], off[: nlev.value + 1]


def art_steps_dev(A: CSRMatrix, d_rows: int, nsteps: int, d_b: int, d_nrA: int, omg: float, d_x: int,
                  ub: float | None = None, lb: float | None = None) -> None:
    """The row steps of basic_sweep (AIRMs.f90:813-858) with device addresses: int32 one-based rows
    (nsteps, each in 1..m), b and nrA (m), x (n, updated in place).  Returns when the sweep is
    done, except with the one-wavefront sweep (FAIR_ART_WINDOW=0, or rows over 1,024 entries), which
    stays queued on the library stream."""
    o = Opts()
    o.relaxpar = float(omg)
    if ub is not None:
        o.set_ubound(ub)
    if lb is not None:
        o.set_lbound(lb)
    co = _opts(o)
    check(lib().fair_art_steps_dev(A.handle, C.c_void_p(d_rows), int(nsteps), C.c_void_p(d_b), C.c_void_p(d_nrA),
                                   C.byref(co), C.c_void_p(d_x)))


STOP_RULES = {"errorgauge": 0, "mutualstep": 1}


def twin_art(art_method: str, stoprule: str, A: CSRMatrix, b, Kin, x0=None, y0=None, order=None,
             slack: int = 7, options: Opts | None = None, uniforms=None, seed: int = 0):
    """call twin_art(xout, final_it, Xstore, Ystore, EG, art_method, stoprule, A, b, Kin, x0, y0,
    order, slack, options), AIRMs.f90:371-811.  Returns (xout, final_it, Xstore, Ystore, EG) with
    Xstore / Ystore of shape (size(Kin), n).  `uniforms`: the random_number draws of randomized
    Kaczmarz in the reference's order (seeded draws when not supplied)."""
    if stoprule not in STOP_RULES:
        raise FairError(FAIR_ERR_METHOD, f"Unknown stopping rule {stoprule}. Use errorgauge or mutualstep.")
    meth, K, ordv = _art_args(art_method, Kin, order)
    b = _f64(b, A.m)
    x0a = None if x0 is None else _f64(x0, A.n)
    y0a = None if y0 is None else _f64(y0, A.n)
    o = _opts(options)
    nord = 0 if ordv is None else ordv.shape[0]
    if uniforms is None and meth == 2:
        cnt = C.c_int64()
        check(lib().fair_twin_art_uniform_bound(A.handle, meth, K.shape[0], _ptr(K), _ptr(ordv), nord, C.byref(o),
                                                C.byref(cnt)))
        uniforms = np.random.default_rng(seed).random(cnt.value)
    u = None if uniforms is None else np.ascontiguousarray(uniforms, np.float64)
    xout = np.zeros(A.n)
    Xs = np.zeros((K.shape[0], A.n))
    Ys = np.zeros((K.shape[0], A.n))
    EG = np.zeros(int(K.max()))
    fi = C.c_int32()
    used = C.c_int64()
    check(lib().fair_twin_art(A.handle, meth, STOP_RULES[stoprule], _ptr(b), K.shape[0], _ptr(K), _ptr(x0a),
                              _ptr(y0a), _ptr(ordv), nord, int(slack), C.byref(o), _ptr(u),
                              0 if u is None else u.shape[0], _ptr(xout), C.byref(fi), _ptr(Xs), _ptr(Ys), _ptr(EG),
                              C.byref(used)))
    return xout, fi.value, Xs, Ys, EG


class SirtGroup:
    """Independent sirt problems (each its own CSRMatrix, sizes may differ) iterated together:
    two launches per iteration for all of them, each iterate bit-identical to its own sirt run.
    The matrices must outlive the group and stay unchanged."""

    def __init__(self, mats, sirt_method: str = "sart"):
        if sirt_method not in SIRT_METHODS:
            raise FairError(FAIR_ERR_METHOD, "Not a recognised SIRT method.")
        self.mats = list(mats)
        arr = (C.c_void_p * len(self.mats))(*[A.handle for A in self.mats])
        h = C.c_void_p()
        check(lib().fair_group_create(len(self.mats), arr, SIRT_METHODS[sirt_method], C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            check(lib().fair_group_free(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self) -> int:
        """Times the grid choices and keeps the fastest: workgroups per CU striding over the work
        items, or 0 for one workgroup per item.  Returns the choice."""
        c = C.c_int32()
        check(lib().fair_group_autotune(self.handle, C.byref(c)))
        return c.value

    def sirt_dev(self, d_bs, maxits: int, d_xs, options: Opts | None = None) -> None:
        """maxits iterations; d_bs / d_xs: device addresses (ints) per problem, x updated in place."""
        if len(d_bs) != len(self.mats) or len(d_xs) != len(self.mats):
            raise FairError(FAIR_ERR_ARG, "one b and one x per problem")
        pb = (C.c_void_p * len(d_bs))(*d_bs)
        px = (C.c_void_p * len(d_xs))(*d_xs)
        o = _opts(options)
        check(lib().fair_group_sirt_dev(self.handle, pb, int(maxits), px, C.byref(o)))

