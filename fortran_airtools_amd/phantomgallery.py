"""module phantomgallery (phantomgallery.f90): the reference's test images, N x N.

Each function returns the image as an (N, N) Fortran-ordered array, img[i, j] = x(i+1, j+1) of
the reference; `vectorise` gives the column-major vector the methods take.  The per-pixel work
(O(N^2 * shapes)) runs on the device through the library; the random parts run here, drawing
from `draws`: the reference's random_number stream in the reference's order (a Draws over the
runtime's draws reproduces the reference's image), or a seeded generator when none is given.
draw_line and swapcoord (:164-249), which nothing in the reference calls, are not provided.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import FairError, FAIR_ERR_ARG, check, lib
from .system_generation import linspace, pi

__all__ = ["Draws", "meshgrid", "intmeshgrid", "choose_phantom", "shepplogan", "smooth", "threephases",
           "threephases_smooth", "binary", "fourphases", "mst", "grains", "bubbles", "vectorise", "randn"]


class Draws:
    """A random_number stream: `uniforms` (the runtime's draws in order) or a seeded generator.
    `consumed` counts the draws taken."""

    def __init__(self, uniforms=None, seed: int = 0):
        self._u = None if uniforms is None else np.ascontiguousarray(uniforms, np.float64).reshape(-1)
        self._g = np.random.default_rng(seed) if uniforms is None else None
        self._buf = np.zeros(0)            # generated, not yet taken (generator streams)
        self.consumed = 0

    def peek(self, k: int) -> np.ndarray:
        """The next k draws (fewer at the end of supplied draws), not taken."""
        if self._u is not None:
            return self._u[self.consumed:self.consumed + k]
        if self._buf.shape[0] < k:
            self._buf = np.concatenate([self._buf, self._g.random(k - self._buf.shape[0])])
        return self._buf[:k]

    def take(self, k: int) -> np.ndarray:
        if self._u is not None and self.consumed + k > self._u.shape[0]:
            raise FairError(FAIR_ERR_ARG, "too few draws for the phantom")
        out = self.peek(k).copy()
        if self._u is None:
            self._buf = self._buf[k:]
        self.consumed += k
        return out

    def one(self) -> float:
        return float(self.take(1)[0])


def _draws(draws):
    return draws if draws is not None else Draws()


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def meshgrid(N: int, corners=None):
    """meshgrid (:17-56) as its two coordinate vectors: X[j] of column j, Y[i] of row i
    (corners = (x0, x1, y0, y1), default the square [-1, 1]^2)."""
    x0, x1, y0, y1 = (-1.0, 1.0, -1.0, 1.0) if corners is None else (float(c) for c in corners)
    return linspace(x0, x1, N), linspace(y1, y0, N)


def intmeshgrid(N: int):
    """intmeshgrid (:58-82): X[j] = j+1, Y[i] = i+1."""
    v = np.arange(1, N + 1, dtype=np.float64)
    return v, v.copy()


def _image(N: int, fn, X, Y, shapes, *extra, x=None):
    out = np.zeros(N * N) if x is None else np.ascontiguousarray(x, np.float64).reshape(-1).copy()
    sh = np.ascontiguousarray(shapes, np.float64).reshape(-1) if len(shapes) else np.zeros(1)
    X = np.ascontiguousarray(X, np.float64)
    Y = np.ascontiguousarray(Y, np.float64)
    nshape = len(shapes)
    check(fn(int(N), _ptr(X), _ptr(Y), int(nshape), _ptr(sh), *extra, _ptr(out)))
    return out


def _as_img(v: np.ndarray, N: int) -> np.ndarray:
    return v.reshape((N, N), order="F")


def vectorise(x) -> np.ndarray:
    """vectorise (:144-162): columns one after another."""
    return np.asarray(x, np.float64).reshape(-1, order="F").copy()


def _norm2_pair(a: float, b: float) -> float:
    """NORM2 of a 2-vector as the reference's compiler evaluates it."""
    ax, ay = abs(a), abs(b)
    mx, mn = (ax, ay) if ax > ay else (ay, ax)
    if mx == 0.0:
        return 0.0
    t = mn / mx
    return mx * math.sqrt(1.0 + t * t)


def _norm2_pairs(dx: np.ndarray, dy: np.ndarray) -> np.ndarray:
    ax, ay = np.abs(dx), np.abs(dy)
    mx, mn = np.where(ax > ay, ax, ay), np.where(ax > ay, ay, ax)
    with np.errstate(invalid="ignore", divide="ignore"):
        t = mn / mx
        d = mx * np.sqrt(1.0 + t * t)
    return np.where(mx == 0.0, 0.0, d)


def randn(draws: Draws | None = None, maxits: int = 100, tolerance: float = 1e-14) -> float:
    """randn (:251-283): one random_number draw u, Newton from 0 on Phi(x) = u."""
    u = _draws(draws).one()
    s2, c2 = math.sqrt(2.0), math.sqrt(2.0 / pi)
    x = 0.0
    for i in range(1, maxits + 1):
        f = 0.5 * (1 + math.erf(x / s2)) - u
        df = c2 * math.exp(-(0.5 * (x * x)))
        if abs(f / df) < tolerance:
            break
        if i == maxits and abs(f / df) > tolerance:
            raise FairError(FAIR_ERR_ARG, "Newton method did not converge.")
        x = x - f / df
    return x


_SHEPP = ((1.0, .6900, .9200, 0.0, 0.0, 0.0), (-.8, .6624, .8740, 0.0, -.0184, 0.0),
          (-.2, .1100, .3100, .22, 0.0, -18.0), (-.2, .1600, .4100, -.22, 0.0, 18.0),
          (.1, .2100, .2500, 0.0, .35, 0.0), (.1, .0460, .0460, 0.0, .1, 0.0),
          (.1, .0460, .0460, 0.0, -.1, 0.0), (.1, .0460, .0230, -.08, -.605, 0.0),
          (.1, .0230, .0230, 0.0, -.606, 0.0), (.1, .0230, .0460, .06, -.605, 0.0))


def shepplogan(N: int) -> np.ndarray:
    """shepplogan (:116-164), the high-contrast Shepp-Logan head."""
    X, Y = meshgrid(N)
    ell = []
    for A, a, b, x0, y0, deg in _SHEPP:
        phi = deg * pi / 180.0
        ell.append((A, a * a, b * b, x0, y0, math.cos(phi), math.sin(phi)))
    return _as_img(_image(N, lib().fair_phantom_ellipses, X, Y, ell), N)


def _f32(v: float) -> float:
    return float(np.float32(v))


def smooth(N: int, p: int = 4) -> np.ndarray:
    """smooth (:166-199): p <= 4 Gaussian bumps, scaled to a maximum of 1."""
    if not 0 <= p <= 4:
        raise FairError(FAIR_ERR_ARG, "smooth has four bumps")
    X, Y = intmeshgrid(N)
    sigma = 0.25 * N
    # rows 2-4 are default-real (single precision) products in the reference
    c = [(0.6 * N, 0.6 * N), (_f32(np.float32(0.5) * np.float32(N)), _f32(np.float32(0.3) * np.float32(N))),
         (_f32(np.float32(0.2) * np.float32(N)), _f32(np.float32(0.7) * np.float32(N))),
         (_f32(np.float32(0.8) * np.float32(N)), _f32(np.float32(0.2) * np.float32(N)))]
    a = (1.0, 0.5, 0.7, 0.9)
    s12 = 1.2 * sigma
    bumps = [(a[i], c[i][0], c[i][1], s12 * s12, sigma * sigma, 1.0) for i in range(p)]
    x = _image(N, lib().fair_phantom_bumps, X, Y, bumps, C.c_int32(2))
    return _as_img(x / x.max(), N)


def _centres(d: Draws, p: int, N: int) -> np.ndarray:
    """call random_number(c) for c(p, 2), then c = c*N: column 1 first."""
    c = d.take(2 * p).reshape((p, 2), order="F")
    return c * N


def _cubic_bumps(N, X, Y, c, den):
    return _image(N, lib().fair_phantom_bumps, X, Y, [(1.0, cx, cy, den, den, 1.0) for cx, cy in c], C.c_int32(3))


def threephases(N: int, p: int = 100, draws: Draws | None = None) -> np.ndarray:
    """threephases (:201-254): two thresholded sums of p cubic bumps, values 0, 1/2, 1."""
    d = _draws(draws)
    X, Y = meshgrid(N, (1.0, float(N), float(N), 1.0))
    s1 = 2.5 * (0.025 * N)
    im1 = _cubic_bumps(N, X, Y, _centres(d, p, N), (s1 * s1) * s1)
    im1 = np.where(im1 < .35, 0.0, 2.0)
    s2 = 2.5 * (0.025 * N)
    im2 = _cubic_bumps(N, X, Y, _centres(d, p, N), s2 * s2)
    im2 = np.where(im2 < 0.55, 0.0, 1.0)
    x = im1 + im2
    x = np.where(x == 3.0, 1.0, x)
    return _as_img(x / x.max(), N)


def threephases_smooth(N: int, p: int = 100, v: float = 1.8, draws: Draws | None = None) -> np.ndarray:
    """threephases_smooth (:256-328): two rescaled thresholded bump sums over a smooth
    background of p Gaussians with random widths."""
    d = _draws(draws)
    X, Y = meshgrid(N, (1.0, float(N), float(N), 1.0))
    s1 = 2.5 * (0.025 * N)
    im1 = _cubic_bumps(N, X, Y, _centres(d, p, N), (s1 * s1) * s1)
    im1 = np.where(im1 < .35, 0.0, im1)
    mn, mx = im1.min(), im1.max()
    im1 = np.where(im1 >= .35, (im1 - mn) / mx * v + .8, im1)
    s2 = 2.5 * (0.025 * N)
    im2 = _cubic_bumps(N, X, Y, _centres(d, p, N), s2 * s2)
    im2 = np.where(im2 < .55, 0.0, im2)
    mn, mx = im2.min(), im2.max()
    im2 = np.where(im2 >= .55, (im2 - mn) / mx * v + .3, im2)
    c = _centres(d, p, N)
    bumps = []
    for i in range(p):
        sb = N * (5.0 * d.one() + .5)
        bumps.append((1e-1 / p, c[i, 0], c[i, 1], sb * sb, sb * sb, 0.5))
    bg = _image(N, lib().fair_phantom_bumps, X, Y, bumps, C.c_int32(2))
    x = np.where(im1 > 0, im1, bg)
    x = np.where(im2 > 0, im2, x)
    return _as_img(x / x.max(), N)


def _walk_disks(N: int, p: int, steps: int, d: Draws, psize: float) -> np.ndarray:
    """The random walks of binary (:358-409): one disk (centre, radius^2) per step taken."""
    disks = []
    for _ in range(p):
        s = d.one()
        vert = d.take(3)
        vert[1] = (10.0 + 10.0 / s) * vert[1]
        vert = np.array([vert[0], vert[0] + vert[1], (vert[0] + vert[1]) + vert[2]])
        vert = vert / vert.max()
        horz = d.take(3)
        horz[1] = 0.0
        horz = np.array([horz[0], horz[0] + horz[1], (horz[0] + horz[1]) + horz[2]])
        horz = horz / horz.max()
        step = .75 * psize * (s + 1)
        pos = 2.0 * d.take(2) - 1.0
        if steps <= 0:
            continue
        # the walk's positions: draws 2i (vertical) and 2i+1 (horizontal) of step i, summed left to
        # right from the start as the reference's loop does; it ends at the first step outside
        pr = d.peek(2 * steps)
        k = pr.shape[0] // 2
        pv, ph = pr[0:2 * k:2], pr[1:2 * k:2]
        dy = np.where(pv < vert[0], -step, np.where(pv > vert[1], step, 0.0))
        dx = np.where(ph < horz[0], -step, np.where(ph > horz[1], step, 0.0))
        py = np.cumsum(np.concatenate([[pos[1]], dy]))[1:]
        px = np.cumsum(np.concatenate([[pos[0]], dx]))[1:]
        out = np.flatnonzero((np.abs(px) > 1.0) | (np.abs(py) > 1.0))
        taken = int(out[0]) if out.size else k
        if taken == k and k < steps:
            d.take(2 * k)                                # the supplied draws end mid-walk
            raise FairError(FAIR_ERR_ARG, "too few draws for the phantom")
        d.take(2 * (taken + 1) if out.size else 2 * k)
        disks.extend(zip(px[:taken], py[:taken], np.full(taken, step * step)))
    return np.array(disks, np.float64).reshape(-1, 3)


def binary(N: int, p: int = 100, steps: int | None = None, draws: Draws | None = None) -> np.ndarray:
    """binary (:330-416): p random walks that stamp a disk at every step; values 0 and 1."""
    d = _draws(draws)
    X, Y = meshgrid(N)
    psize = X[2] - X[1] if N > 2 else float("nan")
    disks = _walk_disks(N, p, N if steps is None else steps, d, psize)
    x = _image(N, lib().fair_phantom_stamp_disks, X, Y, disks, C.c_double(1.0))
    return _as_img(x, N)


def fourphases(N: int, draws: Draws | None = None) -> np.ndarray:
    """fourphases (:418-437): eight binary(N, 80) images with alternating signs, clipped at 0."""
    d = _draws(draws)
    x = np.zeros((N, N), order="F")
    for i in range(1, 9):
        x = x + (-1.0) ** i * binary(N, 80, draws=d)
    x = np.where(x < 0.0, 0.0, x)
    return np.asfortranarray(x / x.max())


def mst(N: int, p: int | None = None, draws: Draws | None = None) -> np.ndarray:
    """mst (:439-527): a minimum spanning tree (Prim, first minimum kept) over p normal points,
    each edge a diamond-shaped band of value 1 + u."""
    d = _draws(draws)
    p = int(30 * math.sqrt(N * 1.0)) if p is None else int(p)
    X, Y = intmeshgrid(N)
    x0 = np.empty((2, p))
    for i in range(2):
        for j in range(p):
            x0[i, j] = .2 * N * randn(d) + .5 * N
    dist = _norm2_pairs(x0[0][:, None] - x0[0][None, :], x0[1][:, None] - x0[1][None, :])
    edges = []
    if p >= 2:
        # best[k]: the smallest distance from loose vertex k to the tree, at the first tree
        # position reaching it; the reference's scan (tree outer, loose inner, strict <) picks
        # the smallest distance, then the earliest tree position, then the earliest loose vertex
        in_tree = np.zeros(p, bool)
        in_tree[0] = True
        pos_of = np.full(p, -1)
        pos_of[0] = 0
        best = dist[0].copy()
        best_from = np.zeros(p, int)
        for t in range(1, p):
            loose = np.flatnonzero(~in_tree)
            bl = best[loose]
            m = bl.min()
            cand = loose[bl == m]
            if cand.size > 1:
                tp = pos_of[best_from[cand]]
                cand = cand[tp == tp.min()]
            ii = int(cand[0])
            jj = int(best_from[ii])
            in_tree[ii] = True
            pos_of[ii] = t
            upd = dist[ii] < best
            best = np.where(upd, dist[ii], best)
            best_from = np.where(upd, ii, best_from)
            xm = .5 * (x0[:, ii] + x0[:, jj])
            th = math.atan2(x0[1, ii] - x0[1, jj], x0[0, ii] - x0[0, jj])
            r = abs(xm[0] - x0[0, jj]) + abs(xm[1] - x0[1, jj])
            u = d.one()
            edges.append((xm[0], xm[1], math.cos(th), math.sin(th), r, 1.0 + u))
    img = _image(N, lib().fair_phantom_last_band, X, Y, edges)
    return _as_img(img / img.max(), N)


def _seed_points(d: Draws, p: int, randomtype: str, scale: float) -> np.ndarray:
    if randomtype == "gaussian":
        x0 = np.empty((2, p))
        for i in range(2):
            for j in range(p):
                v = -2.0
                while not (-1.0 < v < 1.0):
                    v = scale * randn(d)
                x0[i, j] = v
        return x0
    if randomtype == "uniform":
        x0 = d.take(2 * p).reshape((2, p), order="F")
        return 1.5 * (2.0 * x0 - 1.0)
    raise FairError(FAIR_ERR_ARG, "Not a recognised input for randomtype. Use 'gaussian' or 'uniform'.")


def grains(N: int, randomtype: str = "uniform", p: int | None = None, draws: Draws | None = None) -> np.ndarray:
    """grains (:529-597): Voronoi cells of p random seeds, valued by seed index."""
    d = _draws(draws)
    p = int(3 * math.sqrt(N * 1.0)) if p is None else int(p)
    x0 = _seed_points(d, p, randomtype, 1.0)
    d.take(p)                                    # vals(p), drawn and not used
    X, Y = meshgrid(N)
    img = _image(N, lib().fair_phantom_nearest, X, Y, x0.T)
    return _as_img(img / img.max(), N)


def bubbles(N: int, randomtype: str = "uniform", p: int | None = None, draws: Draws | None = None) -> np.ndarray:
    """bubbles (:599-642): a disk at every point, its radius the distance to the nearest other
    point, valued by that neighbour's index; later disks on top."""
    d = _draws(draws)
    p = int(15 * math.sqrt(N * 1.0)) if p is None else int(p)
    X, Y = meshgrid(N)
    x0 = _seed_points(d, p, randomtype, .5)
    dist = _norm2_pairs(x0[0][:, None] - x0[0][None, :], x0[1][:, None] - x0[1][None, :])
    disks = []
    for i in range(p):
        row = dist[i].copy()
        row[i] = np.inf
        kmin = int(np.argmin(row)) if p > 1 else -1      # first of the smallest
        dmin = row[kmin] if p > 1 else np.finfo(np.float64).max
        d.one()                                         # u, drawn and not used
        disks.append((x0[0, i], x0[1, i], (1.0 * dmin) * (1.0 * dmin), float(kmin + 1)))
    img = _image(N, lib().fair_phantom_last_disk, X, Y, disks)
    return _as_img(img / img.max(), N)


def choose_phantom(phantom: str, N: int, draws: Draws | None = None) -> np.ndarray:
    """choose_phantom (:84-114)."""
    table = {"shepplogan": lambda: shepplogan(N), "smooth": lambda: smooth(N),
             "3phases": lambda: threephases(N, draws=draws), "3phasesmooth": lambda: threephases_smooth(N, draws=draws),
             "binary": lambda: binary(N, draws=draws), "4phases": lambda: fourphases(N, draws=draws),
             "mst": lambda: mst(N, draws=draws), "grains": lambda: grains(N, draws=draws),
             "bubbles": lambda: bubbles(N, draws=draws)}
    if phantom not in table:
        raise FairError(FAIR_ERR_ARG, "Not a recognised phantom.")
    return table[phantom]()
