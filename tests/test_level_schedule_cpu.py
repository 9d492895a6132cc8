"""The level schedule of the Kaczmarz sweeps (fair_level_schedule, host work only) on the CPU.

Checked against an independent restatement of its definition below (numpy, one step at a time),
and through what makes a schedule safe to run:
- the steps of one level touch disjoint columns;
- every column sees its steps in sequence order;
- run in schedule order, the oracle's row steps (the reference's basic_sweep loop, restated in C)
  give the sequence order's iterate bit for bit.
A plausible slip fails one of these: a level off by one, an unstable sort, a missed column, a
wrong base for JA.

Matrices come from the oracle's ray tracers:
- the N = 256 fan of the Kaczmarz benchmark (360 x 363 rays), whose cyclic sweep has 1,292 levels;
- a parallel-beam matrix;
- a random pattern with columns outside 1..n, which the sweeps skip.
Sequences: cyclic, symmetric, a random order, repeated rows and pairs, back-and-forth, a single
step."""
import numpy as np
import pytest


def schedule_by_definition(IA, JA, n, seq):
    last = np.zeros(n, np.int64)
    lev = np.empty(len(seq), np.int64)
    for t, r in enumerate(seq):
        cols = JA[IA[r - 1]:IA[r]] - 1
        cols = cols[(cols >= 0) & (cols < n)]
        L = (last[cols].max() if len(cols) else 0) + 1
        last[cols] = L
        lev[t] = L
    order = np.argsort(lev, kind="stable")
    off = np.concatenate([[0], np.cumsum(np.bincount(lev)[1:])])
    return order, off


@pytest.fixture(scope="module")
def fan256(oracle):
    from fortran_airtools_amd import configs as cf
    g = cf.C3
    return oracle.fanlineartomo(g.n, g.p, g.nth, g.r, g.dw, g.sd, g.theta()), g.n * g.n


@pytest.fixture(scope="module")
def par64(oracle):
    return oracle.paralleltomo(64, 91, 60, np.deg2rad(np.linspace(0.0, 179.0, 60))), 64 * 64


def nonempty(IA):
    return np.nonzero(np.diff(IA) > 0)[0].astype(np.int32) + 1


def sequences(IA, rng):
    rows = nonempty(IA)
    return {
        "cyclic": rows,
        "symmetric": np.concatenate([rows, rows[::-1]]),
        "random_order": rng.permutation(rows),
        "repeats": np.repeat(rng.choice(rows, 50), 3),
        "back_and_forth": np.concatenate([rows[100:300], rows[299:99:-1]]),
        "one_step": rows[:1],
    }


def check_schedule(IA, JA, n, seq, rows, off):
    from numpy.testing import assert_array_equal
    order, off_def = schedule_by_definition(IA, JA, n, seq)
    assert_array_equal(off, off_def)
    assert_array_equal(rows, seq[order])
    # disjoint columns within a level
    for l in range(len(off) - 1):
        cols = np.concatenate([JA[IA[r - 1]:IA[r]] for r in rows[off[l]:off[l + 1]]] + [np.zeros(0, JA.dtype)])
        cols = cols[(cols >= 1) & (cols <= n)]
        assert len(np.unique(cols)) == len(cols), f"level {l} touches a column twice"
    # every column sees its steps in sequence order: along the schedule, the positions (in seq)
    # of the steps touching a column increase
    seen = np.full(n + 1, -1, np.int64)
    for pos in order:
        r = seq[pos]
        cols = JA[IA[r - 1]:IA[r]]
        cols = cols[(cols >= 1) & (cols <= n)]
        assert (seen[cols] < pos).all()
        seen[cols] = pos


@pytest.mark.parametrize("which", ["fan256", "par64"])
def test_schedule_matches_definition_and_is_safe(request, which):
    from fortran_airtools_amd import airms
    (IA, JA, _), n = request.getfixturevalue(which)
    rng = np.random.default_rng(21)
    for name, seq in sequences(IA, rng).items():
        if which == "fan256" and name == "random_order":
            seq = seq[:3000]                # keeps the per-step checks short
        rows, off = airms.level_schedule(len(IA) - 1, n, IA, JA, seq)
        check_schedule(IA, JA, n, np.asarray(seq), rows, off)


def test_benchmark_fan_depth(fan256):
    """Depth of the N = 256 fan's sweeps: about p + 5 nth, far below the 56,900 steps."""
    from fortran_airtools_amd import airms
    (IA, JA, _), n = fan256
    rows = nonempty(IA)
    _, off = airms.level_schedule(len(IA) - 1, n, IA, JA, rows)
    assert len(rows) == 56900 and len(off) - 1 == 1292
    _, off2 = airms.level_schedule(len(IA) - 1, n, IA, JA, np.concatenate([rows, rows[::-1]]))
    assert len(off2) - 1 == 2584


def test_columns_outside_the_matrix_are_skipped():
    """Columns outside 1..n (the ray tracers' off-image entries) belong to no level conflict."""
    from fortran_airtools_amd import airms
    rng = np.random.default_rng(5)
    m, n = 200, 300
    lens = rng.integers(0, 12, m)
    IA = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    JA = np.concatenate([rng.choice(np.arange(-3, n + 4), k, replace=False) for k in lens]).astype(np.int32)
    seq = rng.integers(1, m + 1, 500).astype(np.int32)
    rows, off = airms.level_schedule(m, n, IA, JA, seq)
    check_schedule(IA, JA, n, seq, rows, off)


def test_schedule_order_gives_the_same_iterate(oracle, fan256):
    """The oracle's row steps in schedule order equal those in sequence order bit for bit."""
    from fortran_airtools_amd import airms
    (IA, JA, vA), n = fan256
    rng = np.random.default_rng(8)
    rows = nonempty(IA)
    nrA = np.zeros(len(IA) - 1)
    nrA[rows - 1] = np.add.reduceat(vA * vA, IA[rows - 1])
    w = np.where(nrA > 0, nrA, 1.0)
    b = oracle.matvec(IA, JA, vA, rng.random(n))
    x0 = rng.random(n)
    seq = np.concatenate([rows, rows[::-1]])
    sched, _ = airms.level_schedule(len(IA) - 1, n, IA, JA, seq)
    a = oracle.art_steps(IA, JA, vA, b, w, seq, x0, relax=1.2, lb=0.0, ub=1.0)
    c = oracle.art_steps(IA, JA, vA, b, w, sched, x0, relax=1.2, lb=0.0, ub=1.0)
    np.testing.assert_array_equal(a, c)
    # a plausible mistake -- two conflicting steps (adjacent rays sharing pixels) swapped --
    # changes the iterate: the check is not vacuous
    bad = seq.copy()
    bad[[0, 1]] = bad[[1, 0]]
    assert not np.array_equal(oracle.art_steps(IA, JA, vA, b, w, bad, x0, relax=1.2, lb=0.0, ub=1.0), a)


def test_rejects_rows_out_of_range():
    from fortran_airtools_amd import airms
    from fortran_airtools_amd._lib import FairError
    IA = np.array([0, 1, 2], np.int32)
    JA = np.array([1, 2], np.int32)
    with pytest.raises(FairError, match="out of range"):
        airms.level_schedule(2, 2, IA, JA, np.array([1, 3], np.int32))
