"""The phantom gallery and add_white_noise on the device, against images of the compiled reference
(tests/golden/phantoms_small.npz): the Python mirror with the runtime's draws, and the Fortran
drop-in module (fortran_airtools_amd/bin/phantom_check) drawing from its own runtime generator
seeded as the reference was.  Bars as in test_phantoms_cpu: bit for bit without exp / erf (and
threephases, whose exp sums are thresholded), 1e-13 for smooth / threephases_smooth, 1e-12 for
the noise.  At N = 200, where no fixture reaches, the device kernels equal the oracle's numpy
restatement on the same draws (exactly where no exp is involved)."""
import subprocess

import numpy as np
import pytest

from conftest import PHANTOM_CHECK, assert_close_1e12, load_golden, runtime_draws
from test_phantoms_cpu import NumpyKernels, assert_image, case_call, cases

pytestmark = pytest.mark.gpu


def gallery():
    from fortran_airtools_amd import phantomgallery as pg
    return pg


@pytest.mark.parametrize("name", cases())
def test_device_phantom_matches_reference(tmp_path, name):
    pg = gallery()
    g = load_golden("phantoms_small.npz")
    print(name, flush=True)
    got = case_call(pg, g, name, pg.Draws(runtime_draws(int(g[f"{name}/args"][5]), 60000, tmp_path)))
    assert_image(str(g[f"{name}/kind"]), got, g[f"{name}/x"])


@pytest.mark.parametrize("n", [48, 47])
def test_device_shepplogan_matches_reference(n):
    pg = gallery()
    g = load_golden("phantoms_small.npz")
    np.testing.assert_array_equal(pg.vectorise(pg.shepplogan(n)), g[f"shepplogan_{n}/x"])


def run_check(*args):
    r = subprocess.run([PHANTOM_CHECK, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("name", cases())
def test_fortran_phantom_dropin(tmp_path, name):
    g = load_golden("phantoms_small.npz")
    kind = str(g[f"{name}/kind"])
    n, pin, steps, v, gauss, seed = g[f"{name}/args"]
    out = tmp_path / "x.bin"
    print(name, flush=True)
    run_check("phantom", kind, int(n), int(pin), int(steps), float(v), int(gauss), int(seed), out)
    assert_image(kind, np.frombuffer(out.read_bytes(), np.float64), g[f"{name}/x"])


def test_fortran_shepplogan_dropin(tmp_path):
    g = load_golden("phantoms_small.npz")
    out = tmp_path / "x.bin"
    run_check("phantom", "shepplogan", 48, 0, 0, 0.0, 0, 1, out)
    np.testing.assert_array_equal(np.frombuffer(out.read_bytes(), np.float64), g["shepplogan_48/x"])


@pytest.mark.parametrize("k", [0, 1])
def test_white_noise_matches_reference(tmp_path, k):
    from fortran_airtools_amd import system_generation as sg
    g = load_golden("phantoms_small.npz")
    b, eta, seed = g[f"noise{k}/b"], float(g[f"noise{k}/eta"]), int(g[f"noise{k}/seed"])
    bt, e = sg.add_white_noise(b, eta, uniforms=runtime_draws(seed, len(b), tmp_path))
    assert_close_1e12(e, g[f"noise{k}/e"])
    assert_close_1e12(bt, g[f"noise{k}/bt"])
    fin, out = tmp_path / "b.bin", tmp_path / "o.bin"
    with open(fin, "wb") as f:
        np.array([len(b)], np.int32).tofile(f)
        b.astype(np.float64).tofile(f)
    run_check("noise", fin, eta, seed, out)
    res = np.frombuffer(out.read_bytes(), np.float64)
    assert_close_1e12(res[:len(b)], g[f"noise{k}/bt"])
    assert_close_1e12(res[len(b):], g[f"noise{k}/e"])


@pytest.mark.parametrize("kind", ["binary", "bubbles", "grains", "mst", "threephases"])
def test_device_equals_oracle_at_n200(monkeypatch, oracle, kind):
    pg = gallery()
    u = np.random.default_rng(13).random(80000)
    fn = getattr(pg, kind)
    dev = pg.vectorise(fn(200, draws=pg.Draws(u)))
    kern = NumpyKernels()
    monkeypatch.setattr(pg, "lib", lambda: kern)
    ref = pg.vectorise(fn(200, draws=pg.Draws(u)))
    assert_image(kind, dev, ref)
