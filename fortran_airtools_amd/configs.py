"""Workload definitions: geometries, phantoms and noise for the five benchmark problems.

Geometry records follow the reference's `type(data)` (system_generation.f90:6-12): `n` pixels
per side, `p` rays per angle, `nth` angles, and `r`, `dw`, `sd` in units of n (scaled by n at
system_generation.f90:98-100).  Angles are given in degrees and converted as
`theta/180d0*pi`, as `read_data` does (system_generation.f90:526).

Fan parameters (r=2: source at 2n from the centre; sd=4: flat detector 2n beyond the centre;
dw=3.1: the correctly oriented half-fan spans the object's circumscribed circle) are this
project's choice; the reference fixes none.

Phantoms and noise are inputs, generated on the host with numpy from fixed seeds.  The
Shepp-Logan phantom restates phantomgallery.f90:116-164 (with meshgrid :17-56 and
vectorise :672-690); the other phantoms are this project's own seeded generators (the
reference's grains/bubbles draw from the compiler's RANDOM_NUMBER stream and store cell
indices, phantomgallery.f90:577/594).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass(frozen=True)
class Geometry:
    """Mirror of the reference's `type(data)` record (system_generation.f90:6-12)."""

    n: int
    p: int
    nth: int
    r: float = 2.0
    dw: float = 3.1
    sd: float = 4.0
    beam: str = "fan"            # "fan" -> fanlineartomo, "parallel" -> paralleltomo
    deg_start: float = 0.0
    deg_step: float = 1.0

    @property
    def m(self) -> int:
        return self.nth * self.p

    @property
    def ncols(self) -> int:
        return self.n * self.n

    def theta(self) -> np.ndarray:
        deg = self.deg_start + self.deg_step * np.arange(self.nth, dtype=np.float64)
        return deg / 180.0 * math.pi


# C1: plain method, CPU-runnable (N=64, 180 angles 0:2:358 deg, 91 rays).
C1 = Geometry(n=64, p=91, nth=180, deg_step=2.0)
# C2: plain method on GPU at scale (N=512, 720 angles 0:0.5:359.5 deg, 725 rays).
C2 = Geometry(n=512, p=725, nth=720, deg_step=0.5)
# C3: batching + precision (N=256, 360 angles 0:1:359 deg, 363 rays).
C3 = Geometry(n=256, p=363, nth=360, deg_step=1.0)
# C4: matrix-free (N=2048, 2048 angles over 360 deg, 2900 rays).
C4 = Geometry(n=2048, p=2900, nth=2048, deg_step=360.0 / 2048)

# C5: uneven batch -- N drawn from these sizes, angles uniform in N/2..2N over 360 degrees, and
# p the odd number nearest above N/2 (the odd-p fan of the reference, about N/2 detector rays).
C5_SIZES = (128, 192, 256, 384, 512, 768, 1024)


def c5_problems(count: int = 64, seed: int = 5, sizes=None) -> list:
    """The uneven problem set: `count` geometries from one seeded generator (image sizes drawn
    from `sizes`, C5_SIZES by default)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(C5_SIZES if sizes is None else sizes))
        nth = int(rng.integers(n // 2, 2 * n + 1))
        h = n // 2
        p = h + 1 if h % 2 == 0 else h
        out.append(Geometry(n=n, p=p, nth=nth, deg_step=360.0 / nth))
    return out


def linspace(a: float, b: float, n: int) -> np.ndarray:
    """system_generation.f90:17-35: x(i) = c*i + d with c=(b-a)/(n-1), d=a-c."""
    c = (b - a) / (n - 1)
    d = a - c
    return c * np.arange(1, n + 1, dtype=np.float64) + d


_SHEPP = np.array([
    [1.0, .6900, .9200, 0.0, 0.0, 0.0],
    [-.8, .6624, .8740, 0.0, -.0184, 0.0],
    [-.2, .1100, .3100, .22, 0.0, -18.0],
    [-.2, .1600, .4100, -.22, 0.0, 18.0],
    [.1, .2100, .2500, 0.0, .35, 0.0],
    [.1, .0460, .0460, 0.0, .1, 0.0],
    [.1, .0460, .0460, 0.0, -.1, 0.0],
    [.1, .0460, .0230, -.08, -.605, 0.0],
    [.1, .0230, .0230, 0.0, -.606, 0.0],
    [.1, .0230, .0460, .06, -.605, 0.0],
])


def shepplogan(N: int) -> np.ndarray:
    """Vectorised (column-major) Shepp-Logan head, phantomgallery.f90:116-164."""
    lx = linspace(-1.0, 1.0, N)
    ly = linspace(1.0, -1.0, N)
    xc = np.broadcast_to(lx[None, :], (N, N))
    yc = np.broadcast_to(ly[:, None], (N, N))
    img = np.zeros((N, N))
    for A, a, b, x0, y0, phd in _SHEPP:
        phi = phd * math.pi / 180.0
        cp, sp = math.cos(phi), math.sin(phi)
        u = (xc - x0) * cp + (yc - y0) * sp
        v = (yc - y0) * cp - (xc - x0) * sp
        img = np.where((u * u) / (a * a) + (v * v) / (b * b) <= 1.0, img + A, img)
    return np.ascontiguousarray(img.T).reshape(-1)     # column-major vectorise


def _grid(N: int):
    c = (np.arange(N) + 0.5) / N * 2.0 - 1.0
    X, Y = np.meshgrid(c, -c)                            # image rows top->bottom
    return X, Y


def _vec(img: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(img.T).reshape(-1)


def grains(N: int, seeds: int = 200, seed: int = 2) -> np.ndarray:
    """Voronoi cells around `seeds` uniform points of [-1,1]^2, each with a U[0,1] value."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-1.0, 1.0, size=(seeds, 2))
    vals = rng.uniform(0.0, 1.0, size=seeds)
    X, Y = _grid(N)
    best = np.full((N, N), np.inf)
    img = np.zeros((N, N))
    for k in range(seeds):
        d = (X - pts[k, 0]) ** 2 + (Y - pts[k, 1]) ** 2
        sel = d < best
        best[sel] = d[sel]
        img[sel] = vals[k]
    return _vec(img)


def bubbles(N: int, seed: int = 3, count: int = 40) -> np.ndarray:
    """Random discs with U[0.2,1] values on a 0 background, later discs on top."""
    rng = np.random.default_rng(seed)
    X, Y = _grid(N)
    img = np.zeros((N, N))
    for _ in range(count):
        cx, cy = rng.uniform(-0.8, 0.8, 2)
        rad = rng.uniform(0.03, 0.2)
        img[(X - cx) ** 2 + (Y - cy) ** 2 <= rad * rad] = rng.uniform(0.2, 1.0)
    return _vec(img)


def smooth(N: int, seed: int = 3, count: int = 6) -> np.ndarray:
    """Sum of random Gaussians scaled to a maximum of 1."""
    rng = np.random.default_rng(seed)
    X, Y = _grid(N)
    img = np.zeros((N, N))
    for _ in range(count):
        cx, cy = rng.uniform(-0.6, 0.6, 2)
        s = rng.uniform(0.15, 0.45)
        img += rng.uniform(0.3, 1.0) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * s * s))
    return _vec(img / img.max())


def threephases(N: int, seed: int = 3) -> np.ndarray:
    """Three-level thresholded smooth random field (values 0, 0.5, 1)."""
    f = smooth(N, seed=seed, count=12)
    q1, q2 = np.quantile(f, [0.4, 0.75])
    return np.where(f < q1, 0.0, np.where(f < q2, 0.5, 1.0))


def blobs(N: int, seed: int = 4, count: int = 24) -> np.ndarray:
    """Shepp-Logan plus seeded Gaussian blobs, clipped to [0,1]."""
    rng = np.random.default_rng(seed)
    X, Y = _grid(N)
    img = np.zeros((N, N))
    for _ in range(count):
        cx, cy = rng.uniform(-0.6, 0.6, 2)
        s = rng.uniform(0.01, 0.05)
        img += rng.uniform(0.05, 0.2) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * s * s))
    return np.clip(shepplogan(N) + _vec(img), 0.0, 1.0)


PHANTOMS = {"shepplogan": shepplogan, "grains": grains, "bubbles": bubbles, "smooth": smooth,
            "3phases": threephases, "blobs": blobs}


def norm2(v: np.ndarray) -> float:
    """Euclidean norm with numpy's fixed pairwise summation (independent of BLAS threading)."""
    v = np.asarray(v, np.float64)
    return float(np.sqrt(np.sum(v * v)))


def add_noise(b: np.ndarray, eta: float, seed: int) -> np.ndarray:
    """b + e with e = eta*||b||*g/||g||, g standard normal (the rescaling of
    system_generation.f90:742-743; the draws are this project's seeded generator)."""
    g = np.random.default_rng(seed).standard_normal(len(b))
    return b + eta * norm2(b) * g / norm2(g)
