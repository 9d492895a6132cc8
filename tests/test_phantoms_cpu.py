"""The phantom gallery on the CPU, against images of the compiled reference
(tests/golden/phantoms_small.npz, made by tests/golden/make_golden.py --phantoms).

The package's host logic (draw order, walks, spanning tree, seed points, thresholds) runs with
the library's per-pixel kernels replaced by the oracle's numpy restatement (oracle/phantoms.py)
and linspace by the oracle's: this pins both on the CPU.  The runtime draws come from
fortran_airtools_amd/bin/phantom_check, whose seeding is checked against the reference's draws.
Bars: images without exp / erf equal bit for bit; the others (smooth, threephases_smooth) to
1e-13; threephases (thresholded exp sums) equal; noise to 1e-12."""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_close_1e12, load_golden, runtime_draws

EXACT = {"shepplogan", "threephases", "binary", "fourphases", "mst", "grains", "bubbles"}


class NumpyKernels:
    """The fair_phantom_* entry points, restated in numpy (ctypes arguments as the package passes them)."""

    def __init__(self):
        from oracle import phantoms as P
        self.P = P

    @staticmethod
    def _a(ptr, n):
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), shape=(max(n, 1),))[:n]

    def _call(self, fn, n, X, Y, cnt, width, shapes, out, *extra):
        n = int(n)
        o = self._a(out, n * n)
        args = (n, self._a(X, n), self._a(Y, n), self._a(shapes, width * int(cnt)))
        o[:] = fn(*args, *extra, o.copy()) if fn.__name__ in ("last_disk", "stamp_disks", "last_band") else fn(*args, *extra)
        return 0

    def fair_phantom_ellipses(self, n, X, Y, ne, ell, x):
        return self._call(self.P.ellipses, n, X, Y, ne, 7, ell, x)

    def fair_phantom_bumps(self, n, X, Y, nb, bumps, power, x):
        return self._call(self.P.bumps, n, X, Y, nb, 6, bumps, x, power.value)

    def fair_phantom_nearest(self, n, X, Y, npts, pts, x):
        return self._call(self.P.nearest, n, X, Y, npts, 2, pts, x)

    def fair_phantom_last_disk(self, n, X, Y, nd, disks, x):
        return self._call(self.P.last_disk, n, X, Y, nd, 4, disks, x)

    def fair_phantom_stamp_disks(self, n, X, Y, nd, disks, value, x):
        return self._call(self.P.stamp_disks, n, X, Y, nd, 3, disks, x, value.value)

    def fair_phantom_last_band(self, n, X, Y, ne, edges, x):
        return self._call(self.P.last_band, n, X, Y, ne, 6, edges, x)


@pytest.fixture
def gallery(monkeypatch, oracle):
    from fortran_airtools_amd import phantomgallery as pg
    kern = NumpyKernels()
    monkeypatch.setattr(pg, "lib", lambda: kern)
    monkeypatch.setattr(pg, "linspace", oracle.linspace)
    return pg


def cases():
    g = load_golden("phantoms_small.npz")
    return sorted({k.split("/")[0] for k in g.files if k.endswith("/kind")})


def case_call(gallery, g, name, draws):
    """Calls the case's phantom as the reference call of make_golden.py did."""
    kind = str(g[f"{name}/kind"])
    n, pin, steps, v, gauss, _ = g[f"{name}/args"]
    n, pin, steps = int(n), int(pin), int(steps)
    kw = {}
    if pin > 0:
        kw["p"] = pin
    if kind == "binary" and steps > 0:
        kw["steps"] = steps
    if kind == "threephases_smooth" and v > 0:
        kw["v"] = float(v)
    if kind in ("grains", "bubbles") and gauss:
        kw["randomtype"] = "gaussian"
    fn = getattr(gallery, kind)
    return gallery.vectorise(fn(n, **kw) if kind == "smooth" else fn(n, draws=draws, **kw))


def assert_image(kind, got, want):
    if kind in EXACT:
        np.testing.assert_array_equal(got, want)
    else:
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


def test_runtime_draws_match_the_reference(tmp_path):
    g = load_golden("phantoms_small.npz")
    np.testing.assert_array_equal(runtime_draws(7, 64, tmp_path), g["draws_seed7"])


@pytest.mark.parametrize("name", cases())
def test_phantom_matches_reference(gallery, tmp_path, name):
    g = load_golden("phantoms_small.npz")
    seed = int(g[f"{name}/args"][5])
    got = case_call(gallery, g, name, gallery.Draws(runtime_draws(seed, 60000, tmp_path)))
    assert_image(str(g[f"{name}/kind"]), got, g[f"{name}/x"])


@pytest.mark.parametrize("n", [48, 47])
def test_shepplogan_matches_reference(gallery, n):
    g = load_golden("phantoms_small.npz")
    np.testing.assert_array_equal(gallery.vectorise(gallery.shepplogan(n)), g[f"shepplogan_{n}/x"])


@pytest.mark.parametrize("k", [0, 1])
def test_white_noise_newton_matches_reference(tmp_path, k):
    from oracle import phantoms as P
    g = load_golden("phantoms_small.npz")
    b, eta, seed = g[f"noise{k}/b"], float(g[f"noise{k}/eta"]), int(g[f"noise{k}/seed"])
    e = P.normal_newton(runtime_draws(seed, len(b), tmp_path))
    e = eta * np.linalg.norm(b) * e / np.linalg.norm(e)
    assert_close_1e12(e, g[f"noise{k}/e"])
    assert_close_1e12(b + e, g[f"noise{k}/bt"])


def test_choose_phantom_rejects_unknown(gallery):
    from fortran_airtools_amd._lib import FairError
    with pytest.raises(FairError, match="Not a recognised phantom"):
        gallery.choose_phantom("moon", 16)
