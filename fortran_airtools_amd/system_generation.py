"""Python mirror of the reference module `system_generation` (system_generation.f90).

`Data` is the geometry record `type(data)` (:6-12); `fanlineartomo` and `paralleltomo` build
the system matrix on the device (count -> scan -> fill, bit-identical to the reference's
serial two-pass build); `linspace` and `add_white_noise` follow :17-35 and :709-745.
`read_data`, `read_ref` and `read_ordering` read the text files of :494-593 as Fortran's
list-directed input does; `cull_data`, `rescale_to_grid`, `absorbance` and `load_fan_system`
are the steps of the file mode of main.f90:222-275 (b = -log T computed on the device).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import re

import numpy as np

from ._lib import FairError, FAIR_ERR_ARG, Geom, check, lib
from .sparse_matrices import CSRMatrix, read_csr_matrix, store_matrix

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


def normal_newton(u) -> np.ndarray:
    """add_white_noise's inverse normal CDF of uniform draws (system_generation.f90:725-739):
    per draw a Newton solve from 0, on the device."""
    u = np.ascontiguousarray(u, np.float64).reshape(-1)
    e = np.empty_like(u)
    check(lib().fair_normal_newton(int(u.shape[0]), u.ctypes.data_as(C.c_void_p), e.ctypes.data_as(C.c_void_p)))
    return e


def add_white_noise(b, eta: float, g=None, seed: int = 0, uniforms=None):
    """bt = b + e with e = eta*||b||*g/||g|| (system_generation.f90:742-743).  With `uniforms`
    (the runtime's random_number draws, one per sample) g is the reference's Newton inversion of
    them (:725-740, on the device); otherwise g are given or seeded standard normal draws.
    Returns (bt, e)."""
    b = np.asarray(b, np.float64)
    if uniforms is not None:
        u = np.asarray(uniforms, np.float64).reshape(-1)
        if u.shape[0] < b.shape[0]:
            raise FairError(FAIR_ERR_ARG, "one draw per sample is needed")
        g = normal_newton(u[:b.shape[0]])
    elif g is None:
        g = np.random.default_rng(seed).standard_normal(b.shape[0])
    e = eta * np.linalg.norm(b) * g / np.linalg.norm(g)
    return b + e, e


# ---- measured data: the text files of read_data / read_ref / read_ordering ------------------

# One list-directed value: an apostrophe- or quote-delimited string, or a run of characters up to
# a blank or comma (Fortran list-directed input separators).
_VALUE = re.compile(r"'(?:[^']|'')*'|\"(?:[^\"]|\"\")*\"|[^\s,]+")
_REAL = re.compile(r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eEdDqQ][+-]?\d+)?")
_INT = re.compile(r"[+-]?\d+")
_SPECIAL = {"inf", "+inf", "-inf", "infinity", "+infinity", "-infinity", "nan", "+nan", "-nan"}


class _ListDirected:
    """Reads a text file the way Fortran's list-directed READ (fmt=*) does for the statements of
    system_generation.f90:494-593: each READ starts at a new record, takes values from as many
    records as it needs (blank records only separate), and skips the rest of its last record."""

    def __init__(self, path: str):
        try:
            with open(path, "r") as f:
                self.lines = f.read().splitlines()
        except OSError as e:
            raise FairError(FAIR_ERR_ARG, f"cannot open {path}: {e.strerror}") from None
        self.path = path
        self.pos = 0

    def read(self, k: int) -> list[str]:
        out: list[str] = []
        while len(out) < k:
            if self.pos >= len(self.lines):
                raise FairError(FAIR_ERR_ARG, f"end of file in {self.path} (record {self.pos + 1})")
            vals = _VALUE.findall(self.lines[self.pos])
            self.pos += 1
            out.extend(vals[: k - len(out)])
        return out

    def where(self) -> str:
        return f"{self.path}, record {self.pos}"

    def integer(self, tok: str) -> int:
        if not _INT.fullmatch(tok):
            raise FairError(FAIR_ERR_ARG, f"bad integer '{tok}' in {self.where()}")
        return int(tok)

    def real(self, tok: str) -> float:
        if _REAL.fullmatch(tok):
            return float(tok.translate(_EXPONENT))
        if tok.lower() in _SPECIAL:
            return float(tok)
        raise FairError(FAIR_ERR_ARG, f"bad real '{tok}' in {self.where()}")

    def reals(self, count: int) -> np.ndarray:
        """count READs of one real each (the loops of :522-524 and :532-534)."""
        out = np.empty(count)
        for i in range(count):
            out[i] = self.real(self.read(1)[0])
        return out


_EXPONENT = str.maketrans("dDqQ", "eEeE")


def read_data(inputfile: str, verbose: bool = False) -> Data:
    """read_data, system_generation.f90:494-539: six 'label value' records (n, p, nth, r, dw, sd),
    'label l' and l angles in degrees (converted to radians as theta/180*pi, :526), 'label l' and
    l = nth*p transmission values T (p per angle, angle after angle).  The size checks stop with
    the reference's messages (:519, :530); here they raise FairError."""
    if verbose:
        print(" Opening file: ", inputfile.strip())
    f = _ListDirected(inputfile.strip())
    n = f.integer(f.read(2)[1])
    p = f.integer(f.read(2)[1])
    nth = f.integer(f.read(2)[1])
    r = f.real(f.read(2)[1])
    dw = f.real(f.read(2)[1])
    sd = f.real(f.read(2)[1])
    cnt = f.integer(f.read(2)[1])
    if cnt != nth:
        raise FairError(FAIR_ERR_ARG, "Angle vector has the wrong length, nth is incorrect.")
    theta = f.reals(cnt) / 180.0 * pi
    cnt = f.integer(f.read(2)[1])
    if cnt != nth * p:
        raise FairError(FAIR_ERR_ARG, "Data has the wrong size, or p and nth are incorrect.")
    T = f.reals(cnt)
    if verbose:
        print(" Closing file: ", inputfile.strip())
    return Data(n, p, nth, r, dw, sd, theta, T)


def read_ref(inputfile: str, n: int, verbose: bool = False) -> np.ndarray:
    """read_ref, system_generation.f90:541-565: n records of one real (the reference image)."""
    if verbose:
        print(" Opening file: ", inputfile.strip())
    x = _ListDirected(inputfile.strip()).reals(int(n))
    if verbose:
        print(" Closing file: ", inputfile.strip())
    return x


def read_ordering(inputfile: str, verbose: bool = False) -> np.ndarray:
    """read_ordering, system_generation.f90:567-593: 'label l', then l records of one integer
    (a row order, one-based as the reference's art takes it)."""
    if verbose:
        print(" Opening file: ", inputfile.strip())
    f = _ListDirected(inputfile.strip())
    cnt = f.integer(f.read(2)[1])
    order = np.array([f.integer(f.read(1)[0]) for _ in range(max(cnt, 0))], dtype=np.int32)
    if verbose:
        print(" Closing file: ", inputfile.strip())
    return order


def cull_data(dat: Data, skip: int) -> Data:
    """main.f90:229-245: for skip > 0 keep angles 1, 1+skip, 1+2*skip, ... and their p data
    values each (T is (p, nth) in Fortran order); nth becomes the number kept."""
    if skip > 0:
        dat.theta = np.ascontiguousarray(dat._theta()[::skip])
        T = np.asarray(dat.T, np.float64).reshape(dat.nth, dat.p)
        dat.T = np.ascontiguousarray(T[::skip].reshape(-1))
        dat.nth = int(dat.theta.shape[0])
    return dat


def rescale_to_grid(dat: Data) -> float:
    """main.f90:256-259: grid_size = dw*R/sd (the detector width seen at the rotation centre),
    then R, dw and sd in units of grid_size.  Returns grid_size."""
    grid_size = dat.dw * dat.r / dat.sd
    dat.r = dat.r / grid_size
    dat.dw = dat.dw / grid_size
    dat.sd = dat.sd / grid_size
    return grid_size


def absorbance(T, p: int, nth: int, skip: int = 0) -> np.ndarray:
    """b = -log(T) (main.f90:270), after the angle cull of main.f90:229-245 when skip > 0, on the
    device (host arrays in and out)."""
    T = np.ascontiguousarray(T, np.float64)
    if T.shape != (p * nth,):
        raise FairError(FAIR_ERR_ARG, "Data has the wrong size, or p and nth are incorrect.")
    keep = (nth - 1) // skip + 1 if skip > 0 and nth > 0 else nth
    b = np.empty(keep * p)
    check(lib().fair_absorbance(T.ctypes.data_as(C.c_void_p), p, nth, skip, b.ctypes.data_as(C.c_void_p)))
    return b


def absorbance_dev(d_T: int, p: int, nth: int, skip: int, d_b: int) -> None:
    """absorbance with device addresses (asynchronous on the library stream)."""
    check(lib().fair_absorbance_dev(C.c_void_p(d_T), p, nth, skip, C.c_void_p(d_b)))


def load_fan_system(inputfile: str, skip: int = 0, matrixfile: str | None = None, readmatrix: bool = False,
                    storematrix: bool = False, verbose: bool = False):
    """The file mode of main.f90:222-275 as one call: read_data, the angle cull (-c skip), the
    grid_size rescale, A = fanlineartomo scaled by grid_size/N (or read_csr_matrix of
    matrixfile with readmatrix, -r), bt = -log(T), and store_matrix of A with storematrix (-s).
    Returns (A, bt, dat, grid_size); dat holds the culled, rescaled geometry and data."""
    if storematrix and readmatrix:
        raise FairError(FAIR_ERR_ARG, "Choose either storing or reading matrix.")
    if (storematrix or readmatrix) and not matrixfile:
        raise FairError(FAIR_ERR_ARG, "no matrix file named")
    dat = read_data(inputfile, verbose)
    T_read, nth_read = dat.T, dat.nth
    cull_data(dat, skip)
    grid_size = rescale_to_grid(dat)
    if readmatrix:
        A = read_csr_matrix(matrixfile)
    else:
        A = fanlineartomo(dat)
        A.multiply_by_scalar(grid_size / dat.n)
    if A.m != dat.nth * dat.p:
        raise FairError(FAIR_ERR_ARG, "Vector input size incorrect.")
    bt = absorbance(T_read, dat.p, nth_read, skip)
    if storematrix:
        store_matrix(A, matrixfile)
    return A, bt, dat, grid_size
