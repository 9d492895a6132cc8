"""The Kaczmarz row sweep (k_art_sweep) on row sequences and matrices beyond the reference's
fixtures, against the oracle to the 1e-12 bar: the three art variants with box, relaxation,
damping and row probabilities on fan and parallel geometries; explicit row sequences of lengths
around the sweep's unroll, with a row repeated, rows repeated in pairs, adjacent rays and a
forward-then-back order; rows longer than the 512 entries the sweep keeps in registers, and an
image above 65536 pixels."""
import numpy as np
import pytest

from conftest import assert_close_1e12

pytestmark = pytest.mark.gpu


def fair():
    from fortran_airtools_amd import airms, sparse_matrices, system_generation
    return system_generation, sparse_matrices, airms


def parallel(n, p, nth):
    sg, _, _ = fair()
    dat = sg.Data(n=n, p=p, nth=nth, theta=np.deg2rad(np.linspace(0.0, 179.0, nth)))
    return sg.paralleltomo(dat)


def fan(n, p, nth):
    sg, _, _ = fair()
    dat = sg.Data(n=n, p=p, nth=nth, theta=np.deg2rad(np.linspace(0.0, 358.0, nth)))
    return sg.fanlineartomo(dat)


def data(A, seed):
    rng = np.random.default_rng(seed)
    xt = rng.random(A.n)
    return xt, A.matvecmul(xt)


GEOMS = {"parallel_n96": lambda: parallel(96, 137, 60), "fan_n80": lambda: fan(80, 121, 72)}
ART_CASES = [
    ("kaczmarz", dict(relax=1.3, lb=0.0, ub=1.0)),
    ("symkaczmarz", dict(relax=0.9, damp=0.05, lb=0.0)),
    ("randkaczmarz", dict(relax=1.0, rowprob=True, ub=1.0)),
]


@pytest.mark.parametrize("geom", sorted(GEOMS))
@pytest.mark.parametrize("method,o", ART_CASES, ids=[c[0] for c in ART_CASES])
def test_art_variants_match_oracle(oracle, geom, method, o):
    _, _, am = fair()
    A = GEOMS[geom]()
    _, b = data(A, 5)
    opts = am.Opts()
    opts.relaxpar = o["relax"]
    opts.damp = o.get("damp", 0.0)
    opts.rowprobability = o.get("rowprob", False)
    if "lb" in o:
        opts.set_lbound(o["lb"])
    if "ub" in o:
        opts.set_ubound(o["ub"])
    kin = [2, 4] if method == "symkaczmarz" else [1, 3]
    u = None
    if method == "randkaczmarz":
        u = np.random.default_rng(9).random(am.art_uniform_count(method, A, kin, options=opts))
    print(geom, method, A.m, A.n, flush=True)
    got = am.art(method, A, b, kin, options=opts, uniforms=u)
    IA, JA, vA = A.components()
    want, _ = oracle.art(method, IA, JA, vA, A.n, b, kin, u=u, lb=o.get("lb"), ub=o.get("ub"), relax=o["relax"],
                         damp=o.get("damp", 0.0), rowprob=o.get("rowprob", False))
    assert_close_1e12(got, want)


def sequences(m, rng):
    # short and odd lengths; repeated rows; adjacent rays
    out = {f"random_{k}": rng.integers(1, m + 1, k) for k in (1, 2, 3, 4, 5, 6, 59, 60, 61, 64, 119, 120, 121, 301)}
    out["repeat_one_row"] = np.full(37, m // 2)
    out["repeat_pairs"] = np.repeat(rng.integers(1, m + 1, 40), 2)
    out["adjacent_rays"] = np.arange(m // 3, m // 3 + 250) % m + 1
    out["back_and_forth"] = np.concatenate([np.arange(100, 180), np.arange(179, 99, -1)])
    return out


def test_row_sequences_match_oracle(oracle):
    import torch
    from fortran_airtools_amd import _lib
    _, _, am = fair()
    A = parallel(96, 137, 60)
    xt, b = data(A, 7)
    nrA = A.row_norms(2) ** 2
    keep = nrA > 0
    IA, JA, vA = A.components()
    rng = np.random.default_rng(11)
    dev = torch.device("cuda:0")
    d_b = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
    d_w = torch.from_numpy(np.where(keep, nrA, 1.0)).to(dev)
    x0 = rng.random(A.n)
    for name, rows in sequences(A.m, rng).items():
        rows = np.asarray(rows, np.int32)
        rows = rows[keep[rows - 1]]
        d_rows = torch.from_numpy(rows).to(dev)

        d_x = torch.from_numpy(x0.copy()).to(dev)
        print(name, len(rows), flush=True)
        torch.cuda.synchronize()
        am.art_steps_dev(A, d_rows.data_ptr(), len(rows), d_b.data_ptr(), d_w.data_ptr(), 1.1, d_x.data_ptr(),
                         ub=1.0, lb=0.0)
        _lib.check(_lib.lib().fair_synchronize())
        want = oracle.art_steps(IA, JA, vA, b, np.where(keep, nrA, 1.0), rows, x0, relax=1.1, lb=0.0, ub=1.0)
        assert_close_1e12(d_x.cpu().numpy(), want)


def test_long_rows_and_large_images(oracle):
    _, sp, am = fair()
    # a row of 600 entries (n = 2048), and an image of 260 x 260 > 65536 pixels
    rng = np.random.default_rng(3)
    m, n = 300, 2048
    lens = rng.integers(20, 80, m)
    lens[17] = 600
    IA = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    JA = np.concatenate([np.sort(rng.choice(n, k, replace=False)) + 1 for k in lens]).astype(np.int32)
    vA = rng.random(IA[-1])
    big = parallel(260, 180, 8)
    for name, A in (("long_row", sp.CSRMatrix.from_arrays(m, n, IA, JA, vA)), ("n_above_65536", big)):
        _, b = data(A, 8)
        print(name, A.m, A.n, flush=True)
        X = am.art("kaczmarz", A, b, [2], options=am.Opts().set_lbound(0.0))
        I2, J2, v2 = A.components()
        want, _ = oracle.art("kaczmarz", I2, J2, v2, A.n, b, [2], lb=0.0)
        assert_close_1e12(X, want)
