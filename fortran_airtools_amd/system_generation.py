"""Python mirror of the reference module `system_generation` (system_generation.f90).

`Data` is the geometry record `type(data)` (:6-12); `fanlineartomo` and `paralleltomo` build
the system matrix on the device (count -> scan -> fill, bit-identical to the reference's
serial two-pass build); `linspace` and `add_white_noise` follow :17-35 and :709-745.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from ._lib import FairError, FAIR_ERR_ARG, Geom, check, lib
from .sparse_matrices import CSRMatrix

pi = 4.0 * math.atan(1.0)


@dataclasses.dataclass
class Data:
    """type(data): n, p, nth (int); r, dw, sd (units of n); theta (radians); T (data)."""

    n: int
    p: int
    nth: int
    r: float = 2.0
    dw: float = 3.1
    sd: float = 4.0
    theta: np.ndarray | None = None
    T: np.ndarray | None = None

    def _geom(self) -> Geom:
        return Geom(int(self.n), int(self.p), int(self.nth), float(self.r), float(self.dw), float(self.sd))

    def _theta(self) -> np.ndarray:
        if self.theta is None:
            raise FairError(FAIR_ERR_ARG, "theta is not set")
        th = np.ascontiguousarray(self.theta, np.float64)
        if th.shape != (self.nth,):
            raise FairError(FAIR_ERR_ARG, "Angle vector has the wrong length, nth is incorrect.")
        return th


def linspace(a: float, b: float, n: int) -> np.ndarray:
    """system_generation.f90:17-35 (x(i) = c*i + d, evaluated by the library)."""
    x = np.empty(n)
    check(lib().fair_linspace(float(a), float(b), int(n), x.ctypes.data_as(C.c_void_p)))
    return x


def fanlineartomo(dat: Data) -> CSRMatrix:
    """system_generation.f90:79-307."""
    th = dat._theta()
    g = dat._geom()
    h = C.c_void_p()
    check(lib().fair_fanlineartomo(C.byref(g), th.ctypes.data_as(C.c_void_p), C.byref(h)))
    return CSRMatrix(h)


def paralleltomo(dat: Data) -> CSRMatrix:
    """system_generation.f90:309-492."""
    th = dat._theta()
    g = dat._geom()
    h = C.c_void_p()
    check(lib().fair_paralleltomo(C.byref(g), th.ctypes.data_as(C.c_void_p), C.byref(h)))
    return CSRMatrix(h)


def add_white_noise(b, eta: float, g=None, seed: int = 0):
    """bt = b + e with e = eta*||b||*g/||g|| (system_generation.f90:742-743); g standard normal
    draws (the reference inverts the normal CDF of RANDOM_NUMBER draws, :725-740).
    Returns (bt, e)."""
    b = np.asarray(b, np.float64)
    if g is None:
        g = np.random.default_rng(seed).standard_normal(b.shape[0])
    e = eta * np.linalg.norm(b) * g / np.linalg.norm(g)
    return b + e, e
