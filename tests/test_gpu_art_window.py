"""The window sweep of Kaczmarz row sequences without a level schedule (k_win_codes + k_art_window)
against the one-wavefront sweep (k_art_sweep) bit for bit, and against the oracle to the 1e-12
bar.

The window sweep computes each row as the one-wavefront sweep does, on the same values, so the
iterates must be equal, not merely close.  Covered:
- randomized Kaczmarz, uniform order and row-norm sampling, on fan and parallel geometries with
  box and relaxation, one of them with column numbers outside the image;
- twin_art with random orders (both stopping rules);
- basic_sweep on explicit sequences around the window length (32 steps): short ones, a row
  repeated, rows repeated in pairs (every pixel shared), adjacent rays, forward and back, random
  draws with replacement;
- rows above 512 entries (16 entries per lane, windows of 8 steps);
- the N = 256 fan of the Kaczmarz benchmark;
- the device-address entry point;
- the 4 x 4-tile pixel numbering of square images against the image's own order;
- the bounded wait: with a bound of zero polls the call fails with an error and the next call is
  unaffected.
FAIR_DEBUG=1 makes the library print one line per window sweep, which shows that it ran."""
import numpy as np
import pytest

from conftest import assert_close_1e12

pytestmark = pytest.mark.gpu


def fair():
    from fortran_airtools_amd import airms, sparse_matrices, system_generation
    return system_generation, sparse_matrices, airms


def fan(n, p, nth, step=None):
    sg, _, _ = fair()
    theta = np.deg2rad(np.arange(nth) * step) if step else np.deg2rad(np.linspace(0.0, 358.0, nth))
    return sg.fanlineartomo(sg.Data(n=n, p=p, nth=nth, theta=theta))


def parallel(n, p, nth, last_deg=179.0):
    sg, _, _ = fair()
    return sg.paralleltomo(sg.Data(n=n, p=p, nth=nth, theta=np.deg2rad(np.linspace(0.0, last_deg, nth))))


def rhs(A, seed):
    return A.matvecmul(np.random.default_rng(seed).random(A.n))


def window_and_sweep(monkeypatch, capfd, fn):
    """fn() on the window sweep and on the one-wavefront sweep (level schedules off for both);
    returns both results and whether the window sweep printed its line."""
    monkeypatch.setenv("FAIR_ART_LEVELS", "0")
    monkeypatch.setenv("FAIR_DEBUG", "1")
    capfd.readouterr()
    win = fn()
    used = "window sweep" in capfd.readouterr().err
    monkeypatch.setenv("FAIR_ART_WINDOW", "0")
    seq = fn()
    monkeypatch.delenv("FAIR_ART_WINDOW")
    return win, seq, used


def opts(am, relax=1.0, lb=None, ub=None, rowprob=False):
    o = am.Opts()
    o.relaxpar = relax
    o.rowprobability = rowprob
    if lb is not None:
        o.set_lbound(lb)
    if ub is not None:
        o.set_ubound(ub)
    return o


# parallel_n96_pi: the last angle is exactly pi, where the parallel-beam tracer emits 77 column
# numbers above N^2 (a reference quirk kept in the matrix); the sweeps read zero there and store
# nothing.  The reference (and the oracle, which restates it) index x out of bounds there, so that
# case is checked against the one-wavefront sweep only.
GEOMS = {"fan_n80": lambda: fan(80, 121, 72), "parallel_n96": lambda: parallel(96, 137, 60),
         "parallel_n96_pi": lambda: parallel(96, 137, 90, last_deg=180.0)}
CASES = [("uniform", dict(relax=1.2, lb=0.0, ub=1.0)), ("rowprob", dict(relax=1.0, lb=0.0, rowprob=True))]


@pytest.mark.parametrize("geom", sorted(GEOMS))
@pytest.mark.parametrize("name,o", CASES, ids=[c[0] for c in CASES])
def test_randkaczmarz_window_equals_sweep_and_oracle(monkeypatch, capfd, oracle, geom, name, o):
    _, _, am = fair()
    A = GEOMS[geom]()
    if geom.endswith("_pi"):
        assert int((A.components()[1] > A.n).sum()) == 77
    b = rhs(A, 5)
    op = opts(am, **o)
    kin = [1, 3]
    u = np.random.default_rng(9).random(am.art_uniform_count("randkaczmarz", A, kin, options=op))
    print(geom, name, flush=True)
    win, seq, used = window_and_sweep(monkeypatch, capfd,
                                      lambda: am.art("randkaczmarz", A, b, kin, options=op, uniforms=u))
    assert used
    np.testing.assert_array_equal(win, seq)
    if geom.endswith("_pi"):
        return
    IA, JA, vA = A.components()
    want, _ = oracle.art("randkaczmarz", IA, JA, vA, A.n, b, kin, u=u, lb=o.get("lb"), ub=o.get("ub"),
                         relax=o["relax"], rowprob=o.get("rowprob", False))
    assert_close_1e12(win, want)


@pytest.mark.parametrize("rule", ["errorgauge", "mutualstep"])
def test_twin_art_random_window_equals_sweep(monkeypatch, capfd, rule):
    _, _, am = fair()
    A = fan(80, 121, 72)
    b = rhs(A, 7)
    op = opts(am, relax=1.0, lb=0.0, ub=1.0)
    print(rule, flush=True)
    win, seq, used = window_and_sweep(
        monkeypatch, capfd, lambda: am.twin_art("randkaczmarz", rule, A, b, [3, 8], options=op, seed=3))
    assert used
    for a, c in zip(win, seq):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(c))


def test_tiled_numbering_equals_image_order(monkeypatch):
    """For a square image the window sweep runs on the pixels renumbered in 4 x 4 tiles; with
    FAIR_ART_TILED=0 it runs on the image's own order.  Both give the same iterate, bit for bit: a
    90 x 90 image (padded to 92 in the numbering), randomized Kaczmarz with box and relaxation, and
    twin_art's two chains in one launch."""
    _, _, am = fair()
    monkeypatch.setenv("FAIR_ART_LEVELS", "0")
    A = fan(90, 127, 64)
    b = rhs(A, 13)
    op = opts(am, relax=1.1, lb=0.0, ub=1.0)

    def both(fn):
        tiled = fn()
        monkeypatch.setenv("FAIR_ART_TILED", "0")
        plain = fn()
        monkeypatch.delenv("FAIR_ART_TILED")
        return tiled, plain

    t, p = both(lambda: am.art("randkaczmarz", A, b, [1, 3], options=op, seed=2))
    np.testing.assert_array_equal(t, p)
    t, p = both(lambda: am.twin_art("randkaczmarz", "errorgauge", A, b, [4], options=op, seed=5))
    for a, c in zip(t, p):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(c))


def sequences(m, rng):
    out = {f"random_{k}": rng.integers(1, m + 1, k) for k in (2, 3, 16, 31, 32, 33, 63, 64, 65, 301, 2000)}
    out["repeat_one_row"] = np.full(37, m // 2)
    out["repeat_pairs"] = np.repeat(rng.integers(1, m + 1, 40), 2)
    out["adjacent_rays"] = np.arange(m // 3, m // 3 + 250) % m + 1
    out["back_and_forth"] = np.concatenate([np.arange(100, 180), np.arange(179, 99, -1)])
    out["cyclic"] = np.arange(1, m + 1)
    return out


def test_basic_sweep_sequences(monkeypatch, capfd, oracle):
    _, _, am = fair()
    A = parallel(96, 137, 60)
    b = rhs(A, 8)
    nrA = A.row_norms(2) ** 2
    keep = nrA > 0
    w = np.where(keep, nrA, 1.0)
    IA, JA, vA = A.components()
    rng = np.random.default_rng(12)
    x0 = rng.random(A.n)
    for name, rows in sequences(A.m, rng).items():
        rows = np.asarray(rows, np.int32)
        rows = rows[keep[rows - 1]]
        if len(rows) < 2:
            continue
        print(name, len(rows), flush=True)

        def run():
            return am.basic_sweep(x0, A, b, rows, len(rows), False, False, 1.1, w, None, True, True, 1.0, 0.0)

        win, seq, used = window_and_sweep(monkeypatch, capfd, run)
        assert used
        np.testing.assert_array_equal(win, seq)
        want = oracle.art_steps(IA, JA, vA, b, w, rows, x0, relax=1.1, lb=0.0, ub=1.0)
        assert_close_1e12(win, want)


def test_long_rows(monkeypatch, capfd, oracle):
    """Rows of 600 and 1000 entries: 16 entries per lane, windows of 8 steps."""
    _, sp, am = fair()
    rng = np.random.default_rng(3)
    m, n = 300, 2048
    lens = rng.integers(20, 80, m)
    lens[17] = 600
    lens[200] = 1000
    IA = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    JA = np.concatenate([np.sort(rng.choice(n, k, replace=False)) + 1 for k in lens]).astype(np.int32)
    vA = rng.random(IA[-1])
    A = sp.CSRMatrix.from_arrays(m, n, IA, JA, vA)
    b = rhs(A, 9)
    op = opts(am, lb=0.0)
    rows = np.concatenate([np.arange(1, m + 1), np.full(20, 18), np.full(20, 201), rng.integers(1, m + 1, 500)])
    rows = rows.astype(np.int32)
    w = A.row_norms(2) ** 2
    x0 = rng.random(n)

    def run():
        return am.basic_sweep(x0, A, b, rows, len(rows), False, False, 0.9, w, None, False, True, 0.0, 0.0)

    win, seq, used = window_and_sweep(monkeypatch, capfd, run)
    assert used
    np.testing.assert_array_equal(win, seq)
    I2, J2, v2 = A.components()
    want = oracle.art_steps(I2, J2, v2, b, w, rows, x0, relax=0.9, lb=0.0)
    assert_close_1e12(win, want)
    win, seq, used = window_and_sweep(monkeypatch, capfd,
                                      lambda: am.art("randkaczmarz", A, b, [2], options=op, seed=4))
    assert used
    np.testing.assert_array_equal(win, seq)


def test_benchmark_fan_n256(monkeypatch, capfd):
    """N = 256, 360 angles x 363 rays: two sweeps of uniform randomized Kaczmarz."""
    _, _, am = fair()
    A = fan(256, 363, 360, step=1.0)
    b = rhs(A, 10)
    op = opts(am, lb=0.0, ub=1.0)
    win, seq, used = window_and_sweep(monkeypatch, capfd,
                                      lambda: am.art("randkaczmarz", A, b, [2], options=op, seed=6))
    assert used
    np.testing.assert_array_equal(win, seq)


def test_device_entry_point(monkeypatch, capfd):
    import torch
    from fortran_airtools_amd import _lib
    _, _, am = fair()
    A = parallel(96, 137, 60)
    b = rhs(A, 11)
    nrA = A.row_norms(2) ** 2
    rows = np.random.default_rng(13).integers(1, A.m + 1, 900).astype(np.int32)
    rows = rows[nrA[rows - 1] > 0]
    dev = torch.device("cuda:0")
    d_b = torch.from_numpy(b).to(dev)
    d_w = torch.from_numpy(np.where(nrA > 0, nrA, 1.0)).to(dev)
    d_rows = torch.from_numpy(rows).to(dev)
    x0 = np.random.default_rng(14).random(A.n)

    def run():
        d_x = torch.from_numpy(x0.copy()).to(dev)
        torch.cuda.synchronize()
        am.art_steps_dev(A, d_rows.data_ptr(), len(rows), d_b.data_ptr(), d_w.data_ptr(), 1.0, d_x.data_ptr(),
                         ub=1.0, lb=0.0)
        _lib.check(_lib.lib().fair_synchronize())
        return d_x.cpu().numpy()

    win, seq, used = window_and_sweep(monkeypatch, capfd, run)
    assert used
    np.testing.assert_array_equal(win, seq)


def test_bounded_wait_reports_error(monkeypatch):
    """A wait bound of zero polls: the first step that has to wait stops the workgroup, the call
    fails with an error, and the next call (default bound) gives the one-wavefront sweep's result."""
    from fortran_airtools_amd._lib import FairError
    _, _, am = fair()
    A = parallel(96, 137, 60)
    b = rhs(A, 12)
    w = A.row_norms(2) ** 2
    row = int(np.argmax(w)) + 1
    rows = np.full(8, row, np.int32)          # every step depends on the one before
    x0 = np.zeros(A.n)

    def run():
        return am.basic_sweep(x0, A, b, rows, len(rows), False, False, 1.0, w, None, False, False, 0.0, 0.0)

    monkeypatch.setenv("FAIR_ART_LEVELS", "0")
    monkeypatch.setenv("FAIR_ART_WINDOW_SPIN", "0")
    with pytest.raises(FairError, match="waited beyond its bound"):
        run()
    monkeypatch.delenv("FAIR_ART_WINDOW_SPIN")
    got = run()
    monkeypatch.setenv("FAIR_ART_WINDOW", "0")
    np.testing.assert_array_equal(got, run())
