"""CPU restatement (numpy) of the phantom gallery's per-pixel work and of add_white_noise's
per-sample Newton solve -- test infrastructure only: the checker of the device kernels behind
fair_phantom_* / fair_normal_newton, never a path of the product.

Images are column-major n x n (pixel (i, j) at i + j*n); X[j] is the coordinate of column j and
Y[i] that of row i, as meshgrid lays them out (phantomgallery.f90:17-82).  Each function follows
the reference's expression for every shape, in the reference's order, with numpy's IEEE
products and sums:
  ellipses     phantomgallery.f90:151-162 (shepplogan)
  bumps        :182-193 smooth, :224-246 threephases, :287-321 threephases_smooth
  nearest      :580-593 grains (NORM2 of a 2-vector as flang computes it: the larger magnitude
               times sqrt(1 + (smaller/larger)^2), checked against the compiled reference)
  last_disk    :626-636 bubbles
  stamp_disks  :400-407 binary
  last_band    :506-513 mst (pow = 1)
  normal_newton  system_generation.f90:725-739
Pinned by tests/test_phantoms_cpu.py against images of the compiled reference
(tests/golden/phantoms_small.npz): exactly where no exp / erf is involved, to 1e-13 where it is
(numpy's exp is not the host libm's).
"""
from __future__ import annotations

import math

import numpy as np


def _grid(n, X, Y):
    Xg = np.broadcast_to(np.asarray(X, np.float64)[None, :], (n, n))
    Yg = np.broadcast_to(np.asarray(Y, np.float64)[:, None], (n, n))
    return Xg, Yg


def _flat(img):
    return np.ascontiguousarray(img.T).reshape(-1)   # column-major


def _img(x, n):
    return np.asarray(x, np.float64).reshape((n, n), order="F").copy()


def norm2_pair(dx, dy):
    ax, ay = np.abs(dx), np.abs(dy)
    mx, mn = np.where(ax > ay, ax, ay), np.where(ax > ay, ay, ax)
    with np.errstate(invalid="ignore", divide="ignore"):
        t = mn / mx
        d = mx * np.sqrt(1.0 + t * t)
    return np.where(mx == 0.0, 0.0, d)


def ellipses(n, X, Y, ell):
    Xg, Yg = _grid(n, X, Y)
    v = np.zeros((n, n))
    for A, a2, b2, x0, y0, c, s in np.asarray(ell, np.float64).reshape(-1, 7):
        dx, dy = Xg - x0, Yg - y0
        u = dx * c + dy * s
        w = dy * c - dx * s
        v = np.where((u * u) / a2 + (w * w) / b2 <= 1.0, v + A, v)
    return _flat(v)


def bumps(n, X, Y, bumps_, power):
    Xg, Yg = _grid(n, X, Y)
    v = np.zeros((n, n))

    def P(t):
        if power == 2:
            return t * t
        a = np.abs(t)
        return (a * a) * a

    for w, cx, cy, dx, dy, c in np.asarray(bumps_, np.float64).reshape(-1, 6):
        tx = (c * P(Xg - cx)) / dx
        ty = (c * P(Yg - cy)) / dy
        v = v + w * np.exp(-tx - ty)
    return _flat(v)


def nearest(n, X, Y, pts):
    Xg, Yg = _grid(n, X, Y)
    dmin = np.full((n, n), np.inf)
    kmin = np.zeros((n, n))
    for q, (px, py) in enumerate(np.asarray(pts, np.float64).reshape(-1, 2)):
        d = norm2_pair(Xg - px, Yg - py)
        sel = d < dmin
        dmin = np.where(sel, d, dmin)
        kmin = np.where(sel, q + 1.0, kmin)
    return _flat(kmin)


def last_disk(n, X, Y, disks, x):
    Xg, Yg = _grid(n, X, Y)
    v = _img(x, n)
    for cx, cy, r2, val in np.asarray(disks, np.float64).reshape(-1, 4):
        dx, dy = Xg - cx, Yg - cy
        v = np.where(dx * dx + dy * dy <= r2, val, v)
    return _flat(v)


def stamp_disks(n, X, Y, disks, value, x):
    Xg, Yg = _grid(n, X, Y)
    v = _img(x, n)
    for cx, cy, r2 in np.asarray(disks, np.float64).reshape(-1, 3):
        dx, dy = Xg - cx, Yg - cy
        v = np.where(dx * dx + dy * dy <= r2, value, v)
    return _flat(v)


def last_band(n, X, Y, edges, x):
    Xg, Yg = _grid(n, X, Y)
    v = _img(x, n)
    for m1, m2, c, s, r, val in np.asarray(edges, np.float64).reshape(-1, 6):
        tx, ty = Xg - m1, Yg - m2
        u = c * tx + s * ty
        w = -(s * tx) + c * ty
        v = np.where(np.abs(u) + np.abs(w) < r, val, v)
    return _flat(v)


def normal_newton(u):
    s2, c2 = math.sqrt(2.0), math.sqrt(2.0 / (4.0 * math.atan(1.0)))
    out = np.empty(len(u))
    for k, uk in enumerate(np.asarray(u, np.float64)):
        x = 0.0
        for i in range(1, 101):
            f = 0.5 * (1.0 + math.erf(x / s2)) - uk
            df = c2 * math.exp(-(0.5 * (x * x)))
            r = f / df
            if abs(r) < 1e-14:
                break
            if i == 100 and abs(r) > 1e-14:
                raise ArithmeticError("Newton method did not converge.")
            x = x - r
        out[k] = x
    return out
