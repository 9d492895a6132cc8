"""The level-parallel Kaczmarz sweeps (k_art_level, one launch per level; k_art_levels, one
workgroup) against the one-wavefront sweep (k_art_sweep) bit for bit, and against the oracle to
the 1e-12 bar.

A level schedule sorts a row sequence so that the steps of one level touch disjoint pixels; every
pixel then sees its updates in sequence order and each row is computed as the one-wavefront
sweep computes it, so the iterates must be equal, not merely close.  Covered: art (kaczmarz,
symkaczmarz, a caller's order) and twin_art (both stopping rules) on fan and parallel geometries
with box, relaxation and damping; basic_sweep on explicit sequences (short, repeated rows, pairs,
back-and-forth, random) with the width threshold lowered to 1 so that narrow schedules run on
the level path too; rows above the 512 entries a wavefront keeps in registers; the N = 256 fan
of the Kaczmarz benchmark (about 44 steps per level) and the N = 512 SART problem (85 steps per
level, rows up to about 1000 entries).  Every test runs with both executors of a
schedule: one workgroup of 16 wavefronts with a barrier between levels, and one launch per level
captured into a graph (FAIR_ART_LEVEL_MODE=graph).  FAIR_DEBUG=1 makes the library print one line
per level sweep, which shows that the level path ran."""
import numpy as np
import pytest

from conftest import assert_close_1e12

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["workgroup", "graph"])
def level_mode(request, monkeypatch):
    """Both level executors: one workgroup walking the levels, and one launch per level (graph)."""
    monkeypatch.setenv("FAIR_ART_LEVEL_MODE", request.param)
    return request.param


def fair():
    from fortran_airtools_amd import airms, sparse_matrices, system_generation
    return system_generation, sparse_matrices, airms


def fan(n, p, nth, step=None):
    sg, _, _ = fair()
    theta = np.deg2rad(np.arange(nth) * step) if step else np.deg2rad(np.linspace(0.0, 358.0, nth))
    return sg.fanlineartomo(sg.Data(n=n, p=p, nth=nth, theta=theta))


def parallel(n, p, nth):
    sg, _, _ = fair()
    return sg.paralleltomo(sg.Data(n=n, p=p, nth=nth, theta=np.deg2rad(np.linspace(0.0, 179.0, nth))))


def rhs(A, seed):
    return A.matvecmul(np.random.default_rng(seed).random(A.n))


def both_paths(monkeypatch, capfd, fn, width="1"):
    """fn() with the level sweep (FAIR_ART_LEVEL_WIDTH=width) and with the one-wavefront sweep
    (FAIR_ART_LEVELS=0, FAIR_ART_WINDOW=0);
    returns both results and whether the level path printed its line."""
    monkeypatch.setenv("FAIR_DEBUG", "1")
    monkeypatch.setenv("FAIR_ART_LEVEL_WIDTH", width)
    capfd.readouterr()
    lev = fn()
    used = "level sweep" in capfd.readouterr().err
    monkeypatch.setenv("FAIR_ART_LEVELS", "0")
    monkeypatch.setenv("FAIR_ART_WINDOW", "0")
    seq = fn()
    monkeypatch.delenv("FAIR_ART_LEVELS")
    monkeypatch.delenv("FAIR_ART_WINDOW")
    return lev, seq, used


def opts(am, relax=1.0, damp=0.0, lb=None, ub=None):
    o = am.Opts()
    o.relaxpar = relax
    o.damp = damp
    if lb is not None:
        o.set_lbound(lb)
    if ub is not None:
        o.set_ubound(ub)
    return o


GEOMS = {"fan_n80": lambda: fan(80, 121, 72), "parallel_n96": lambda: parallel(96, 137, 60)}
CASES = [("kaczmarz", dict(relax=1.3, lb=0.0, ub=1.0), [1, 3]),
         ("symkaczmarz", dict(relax=0.9, damp=0.05, lb=0.0), [2, 4])]


@pytest.mark.parametrize("geom", sorted(GEOMS))
@pytest.mark.parametrize("method,o,kin", CASES, ids=[c[0] for c in CASES])
def test_art_levels_equal_sweep_and_oracle(monkeypatch, capfd, oracle, geom, method, o, kin):
    _, _, am = fair()
    A = GEOMS[geom]()
    b = rhs(A, 5)
    op = opts(am, **o)
    print(geom, method, flush=True)
    lev, seq, used = both_paths(monkeypatch, capfd, lambda: am.art(method, A, b, kin, options=op))
    assert used
    np.testing.assert_array_equal(lev, seq)
    IA, JA, vA = A.components()
    want, _ = oracle.art(method, IA, JA, vA, A.n, b, kin, lb=o.get("lb"), ub=o.get("ub"), relax=o["relax"],
                         damp=o.get("damp", 0.0))
    assert_close_1e12(lev, want)


def test_art_caller_order_and_repeat_call(monkeypatch, capfd):
    _, _, am = fair()
    A = fan(80, 121, 72)
    b = rhs(A, 6)
    order = np.random.default_rng(4).permutation(A.m)[: A.m // 2] + 1
    op = opts(am, relax=1.1, lb=0.0)
    lev, seq, used = both_paths(monkeypatch, capfd, lambda: am.art("kaczmarz", A, b, [2], order=order, options=op))
    assert used
    np.testing.assert_array_equal(lev, seq)
    # a second call reuses the cached schedule and gives the same iterate
    monkeypatch.setenv("FAIR_ART_LEVEL_WIDTH", "1")
    np.testing.assert_array_equal(am.art("kaczmarz", A, b, [2], order=order, options=op), lev)


@pytest.mark.parametrize("rule", ["errorgauge", "mutualstep"])
@pytest.mark.parametrize("method", ["kaczmarz", "symkaczmarz"])
def test_twin_art_levels_equal_sweep(monkeypatch, capfd, rule, method):
    _, _, am = fair()
    A = fan(80, 121, 72)
    b = rhs(A, 7)
    kin = [2, 6] if method == "symkaczmarz" else [3, 8]
    op = opts(am, relax=1.0, lb=0.0, ub=1.0)
    print(method, rule, flush=True)
    lev, seq, used = both_paths(monkeypatch, capfd, lambda: am.twin_art(method, rule, A, b, kin, options=op))
    assert used
    for a, c in zip(lev, seq):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(c))


@pytest.mark.parametrize("order", [[90], [80, 81, 80], [70, 71, 72, 71, 70]], ids=["one_row", "aba", "abcba"])
def test_twin_art_palindromic_order(monkeypatch, capfd, order):
    """A palindromic order gives both chains the same sequence, hence one cached schedule."""
    _, _, am = fair()
    A = fan(80, 121, 72)
    b = rhs(A, 11)
    op = opts(am, relax=1.0, lb=0.0)
    rows = np.asarray(order, np.int32)
    rows = rows[A.row_norms(2)[rows - 1] > 0]
    assert len(rows) == len(order)
    lev, seq, used = both_paths(monkeypatch, capfd,
                                lambda: am.twin_art("kaczmarz", "errorgauge", A, b, [3, 6], order=rows, options=op))
    assert used
    for a, c in zip(lev, seq):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(c))


def sequences(m, rng):
    out = {f"random_{k}": rng.integers(1, m + 1, k) for k in (1, 2, 15, 16, 17, 33, 301, 2000)}
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
        if len(rows) == 0:
            continue
        print(name, len(rows), flush=True)

        def run():
            return am.basic_sweep(x0, A, b, rows, len(rows), False, False, 1.1, w, None, True, True, 1.0, 0.0)

        lev, seq, used = both_paths(monkeypatch, capfd, run)
        assert used
        np.testing.assert_array_equal(lev, seq)
        want = oracle.art_steps(IA, JA, vA, b, w, rows, x0, relax=1.1, lb=0.0, ub=1.0)
        assert_close_1e12(lev, want)


def test_long_rows_and_large_image(monkeypatch, capfd, oracle):
    _, sp, am = fair()
    rng = np.random.default_rng(3)
    m, n = 300, 2048
    lens = rng.integers(20, 80, m)
    lens[17] = 600
    lens[200] = 1500
    IA = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    JA = np.concatenate([np.sort(rng.choice(n, k, replace=False)) + 1 for k in lens]).astype(np.int32)
    vA = rng.random(IA[-1])
    for name, A in (("long_rows", sp.CSRMatrix.from_arrays(m, n, IA, JA, vA)), ("n_above_65536", parallel(260, 180, 8))):
        b = rhs(A, 9)
        op = opts(am, lb=0.0)
        print(name, flush=True)
        lev, seq, used = both_paths(monkeypatch, capfd, lambda: am.art("kaczmarz", A, b, [2], options=op))
        assert used
        np.testing.assert_array_equal(lev, seq)
        I2, J2, v2 = A.components()
        want, _ = oracle.art("kaczmarz", I2, J2, v2, A.n, b, [2], lb=0.0)
        assert_close_1e12(lev, want)


def test_benchmark_fan_default_width(monkeypatch, capfd):
    """N = 256, 360 angles x 363 rays: the default threshold picks the level path by itself."""
    _, _, am = fair()
    A = fan(256, 363, 360, step=1.0)
    b = rhs(A, 10)
    op = opts(am, lb=0.0, ub=1.0)
    lev, seq, used = both_paths(monkeypatch, capfd, lambda: am.art("symkaczmarz", A, b, [2], options=op), width="4")
    assert used
    np.testing.assert_array_equal(lev, seq)


def test_sart_problem_n512(monkeypatch, capfd, oracle):
    """N = 512, 720 x 725 rays (rows up to about 1000 entries, levels built with 16 entries per
    lane): two cyclic sweeps, bit for bit against the one-wavefront sweep and to 1e-12 against the
    oracle."""
    _, _, am = fair()
    from fortran_airtools_amd import configs as cf
    sg, _, _ = fair()
    g = cf.C2
    A = sg.fanlineartomo(sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta()))
    b = rhs(A, 12)
    op = opts(am, relax=1.0, lb=0.0, ub=1.0)
    lev, seq, used = both_paths(monkeypatch, capfd, lambda: am.art("kaczmarz", A, b, [2], options=op), width="4")
    assert used
    np.testing.assert_array_equal(lev, seq)
    IA, JA, vA = A.components()
    want, _ = oracle.art("kaczmarz", IA, JA, vA, A.n, b, [2], lb=0.0, ub=1.0)
    assert_close_1e12(lev, want)


def test_schedule_cache_eviction(monkeypatch, capfd):
    """More sequences than the matrix caches (six): the oldest schedules and their graphs are
    dropped and planned again when reused; every result equals the one-wavefront sweep's."""
    _, _, am = fair()
    A = parallel(96, 137, 60)
    b = rhs(A, 13)
    nrA = A.row_norms(2) ** 2
    rows = np.nonzero(nrA > 0)[0].astype(np.int32) + 1
    w = np.where(nrA > 0, nrA, 1.0)
    x0 = np.random.default_rng(14).random(A.n)
    seqs = [np.roll(rows, 37 * k) for k in range(8)] + [rows]
    for k, seq in enumerate(seqs):
        print("sequence", k, flush=True)

        def run():
            return am.basic_sweep(x0, A, b, seq, len(seq), False, False, 1.0, w, None, True, True, 1.0, 0.0)

        lev, ref, used = both_paths(monkeypatch, capfd, run)
        assert used
        np.testing.assert_array_equal(lev, ref)


def test_twin_art_after_cache_eviction(monkeypatch, capfd):
    """A schedule in use stays alive while a later plan of the same call evicts entries: art plans
    the order's sequence, five basic_sweep calls fill the six-entry cache behind it, then twin_art
    with that order hits the cached forward sequence and plans the reversed one (a miss with the
    cache full).  The forward schedule must survive, and the twin result equals the one-wavefront
    sweep's bit for bit."""
    _, _, am = fair()
    A = fan(80, 121, 72)
    b = rhs(A, 21)
    nrA = A.row_norms(2) ** 2
    rows = np.nonzero(nrA > 0)[0].astype(np.int32) + 1
    w = np.where(nrA > 0, nrA, 1.0)
    order = np.random.default_rng(22).permutation(rows)[: len(rows) // 2].astype(np.int32)
    x0 = np.random.default_rng(23).random(A.n)
    op = opts(am, relax=1.0, lb=0.0, ub=1.0)

    def run():
        am.art("kaczmarz", A, b, [1], order=order, options=op)
        for k in range(5):
            am.basic_sweep(x0, A, b, np.roll(rows, 53 * (k + 1)), len(rows), False, False, 1.0, w, None, True, True,
                           1.0, 0.0)
        return am.twin_art("kaczmarz", "errorgauge", A, b, [2, 4], order=order, options=op)

    lev, seq, used = both_paths(monkeypatch, capfd, run)
    assert used
    for a, c in zip(lev, seq):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(c))


def test_duplicate_columns_refused_by_kaczmarz():
    """A caller-made matrix whose row holds a column twice: products sum the entries as the
    reference's loops do, while the Kaczmarz paths (whose lanes update every entry of a row from
    the same x) refuse it with an error instead of a silently different iterate."""
    from fortran_airtools_amd._lib import FairError
    _, sm, am = fair()
    IA = np.array([0, 3, 5], np.int32)
    JA = np.array([1, 3, 1, 2, 3], np.int32)
    vA = np.array([1.0, 2.0, 0.5, 4.0, 1.5])
    A = sm.CSRMatrix.from_arrays(2, 3, IA, JA, vA)
    x = np.array([1.0, 2.0, 3.0])
    np.testing.assert_array_equal(A.matvecmul(x), [1.0 + 2.0 * 3.0 + 0.5, 4.0 * 2.0 + 1.5 * 3.0])
    b = np.array([1.0, 1.0])
    with pytest.raises(FairError, match="same column twice"):
        am.art("kaczmarz", A, b, [1])
    with pytest.raises(FairError, match="same column twice"):
        am.twin_art("kaczmarz", "errorgauge", A, b, [2])
