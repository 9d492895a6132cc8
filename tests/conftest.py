"""Shared fixtures.  GPU tests carry @pytest.mark.gpu; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfair_hip.so)")


def load_golden(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name} not present")
    return np.load(path, allow_pickle=False)


def small_cases():
    g = load_golden("sysgen_small.npz")
    names = sorted({k.split("/")[0] for k in g.files})
    out = []
    for nm in names:
        n, p, nth = (int(v) for v in g[f"{nm}/geom"])
        r, dw, sd = (float(v) for v in g[f"{nm}/rdwsd"])
        out.append(dict(name=nm, beam=str(g[f"{nm}/beam"]), n=n, p=p, nth=nth, r=r, dw=dw, sd=sd,
                        theta=g[f"{nm}/theta"], IA=g[f"{nm}/IA"], JA=g[f"{nm}/JA"], vA=g[f"{nm}/vA"]))
    return out


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle, build_oracle
    build_oracle()
    return Oracle()


def bitequal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def relerr(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return np.linalg.norm(x - ref) / (den if den > 0 else 1.0)


def assert_close_1e12(x, ref, tol=1e-12):
    """The plain-method bar: norm-wise relative error and max-abs error both within tol."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    assert x.shape == ref.shape
    assert np.all(np.isfinite(x) == np.isfinite(ref))
    fin = np.isfinite(ref)
    r = relerr(x[fin], ref[fin])
    mabs = np.max(np.abs(x[fin] - ref[fin])) if fin.any() else 0.0
    scale = max(np.max(np.abs(ref[fin])) if fin.any() else 0.0, 1e-300)
    assert r <= tol, f"relative error {r:.3e} > {tol}"
    assert mabs <= tol * scale, f"max-abs error {mabs:.3e} > {tol}*{scale:.3e}"


PHANTOM_CHECK = os.path.join(ROOT, "fortran_airtools_amd", "bin", "phantom_check")


def runtime_draws(seed, count, tmp_path):
    """`count` random_number draws of the runtime seeded with put(seed + 7919*i), written by
    fortran_airtools_amd/bin/phantom_check (the seeding the golden fixtures were made with)."""
    import subprocess
    if not os.path.exists(PHANTOM_CHECK):
        pytest.skip("fortran_airtools_amd/bin/phantom_check is not built")
    out = tmp_path / f"draws_{seed}_{count}.bin"
    r = subprocess.run([PHANTOM_CHECK, "draws", str(seed), str(count), str(out)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.frombuffer(out.read_bytes(), np.float64)

