"""Matrix-free fan-beam projector (FP32) over the C ABI.

No counterpart in the reference: its README lists "make matrix-free implementations" as a to-do
(README.md:7).  The operator is the one `fanlineartomo` stores (system_generation.f90:79-307),
with the intersections recomputed on every product, so a 2048^2 image with 2048 x 2900 rays
(about 1.1e10 ray-pixel segments, beyond the reference's int32 nnz) can be reconstructed.
Products take the reference's vector layouts: image n*n in column-major pixel order, sinogram
nth*p angle-major.  `normest` and `gradient_descent` keep the reference's control flow
(AIRMs.f90:1129-1173, 860-995) with FP32 vectors and FP64 norms.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import FairError, FAIR_ERR_ARG, check, lib
from .airms import Opts, _opts
from .system_generation import Data


def _f32(a, n):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 1 or a.shape[0] != n:
        raise FairError(FAIR_ERR_ARG, "Vector input size incorrect.")
    return a


class FanProjector:
    """A of fanlineartomo(dat), applied without storing it."""

    def __init__(self, dat: Data):
        th = dat._theta()
        g = dat._geom()
        h = C.c_void_p()
        check(lib().fair_mf_create(C.byref(g), th.ctypes.data_as(C.c_void_p), C.byref(h)))
        self.handle = h
        self.m = int(dat.nth) * int(dat.p)
        self.n = int(dat.n) ** 2

    def close(self):
        if self.handle:
            check(lib().fair_mf_free(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def segments(self) -> int:
        """Ray-pixel segments of positive length in one product (the operator's nonzeros)."""
        c = C.c_int64()
        check(lib().fair_mf_segments(self.handle, C.byref(c)))
        return c.value

    # ---- host-array products ---------------------------------------------------------------
    def matvecmul(self, x) -> np.ndarray:
        x = _f32(x, self.n)
        y = np.empty(self.m, np.float32)
        check(lib().fair_mf_forward(self.handle, x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p)))
        return y

    def transvecmul(self, y) -> np.ndarray:
        y = _f32(y, self.m)
        x = np.empty(self.n, np.float32)
        check(lib().fair_mf_adjoint(self.handle, y.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p)))
        return x

    # ---- device products (HBM pointers as integers) ---------------------------------------
    def matvecmul_dev(self, d_x: int, d_y: int) -> None:
        check(lib().fair_mf_forward_dev(self.handle, C.c_void_p(d_x), C.c_void_p(d_y)))

    def transvecmul_dev(self, d_y: int, d_x: int) -> None:
        check(lib().fair_mf_adjoint_dev(self.handle, C.c_void_p(d_y), C.c_void_p(d_x)))

    def residual_dev(self, d_x: int, d_b: int, d_r: int) -> None:
        check(lib().fair_mf_residual_dev(self.handle, C.c_void_p(d_x), C.c_void_p(d_b), C.c_void_p(d_r)))

    # ---- methods ------------------------------------------------------------------------------
    def normest(self, tolerance: float = 1e-5, max_its: int = 100, v0=None, seed: int = 0) -> float:
        v0 = np.random.default_rng(seed).random(self.n) if v0 is None else v0
        v0 = _f32(v0, self.n)
        s = C.c_double()
        check(lib().fair_mf_normest(self.handle, v0.ctypes.data_as(C.c_void_p), float(tolerance), int(max_its),
                                    C.byref(s)))
        return s.value

    def gradient_descent_dev(self, d_b: int, d_x: int, max_iter: int = 100, rho_in: float = 1.5,
                             tolerance: float = 1e-6, sigma: float | None = None,
                             options: Opts | None = None) -> int:
        """In-place on the device iterate d_x (x0 in, result out); returns accepted updates."""
        if tolerance < 0:
            raise FairError(FAIR_ERR_ARG, "Must use positive tolerance.")
        if sigma is None:
            sigma = self.normest()
        its = C.c_int32()
        o = _opts(options)
        check(lib().fair_mf_gd_dev(self.handle, C.c_void_p(d_b), int(max_iter), float(rho_in), float(tolerance),
                                   float(sigma), C.byref(o), C.c_void_p(d_x), C.byref(its)))
        return its.value
