"""The oracle (C restatement, oracle/fair_oracle.c) against the reference's own golden vectors.

Golden vectors come from the reference's Fortran compiled with flang (tests/golden/
make_golden.py).  Everything here is bit-exact except normest, whose final division differs
from the reference's NORM2/DOT_PRODUCT in the last bits.
"""
import hashlib

import numpy as np
import pytest

from conftest import bitequal, load_golden, small_cases


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("case", small_cases(), ids=lambda c: c["name"])
def test_small_geometries_bitexact(oracle, case):
    if case["beam"] == "fan":
        IA, JA, vA = oracle.fanlineartomo(case["n"], case["p"], case["nth"], case["r"], case["dw"],
                                          case["sd"], case["theta"])
    else:
        IA, JA, vA = oracle.paralleltomo(case["n"], case["p"], case["nth"], case["theta"])
    assert bitequal(IA, case["IA"]) and bitequal(JA, case["JA"]) and bitequal(vA, case["vA"])


@pytest.fixture(scope="module")
def c1(oracle):
    from fortran_airtools_amd import configs as cf
    g = cf.C1
    IA, JA, vA = oracle.fanlineartomo(g.n, g.p, g.nth, g.r, g.dw, g.sd, g.theta())
    return load_golden("c1.npz"), IA, JA, vA, g.ncols


def test_c1_matrix(c1):
    gold, IA, JA, vA, n = c1
    assert bitequal(IA, gold["IA"])
    assert sha(JA) == str(gold["JA_sha"]) and sha(vA) == str(gold["vA_sha"])
    rows = gold["sample_rows"]
    sel = np.concatenate([np.arange(IA[k], IA[k + 1]) for k in rows])
    assert bitequal(JA[sel], gold["sample_JA"]) and bitequal(vA[sel], gold["sample_vA"])


def test_c1_products(oracle, c1):
    gold, IA, JA, vA, n = c1
    assert bitequal(oracle.matvec(IA, JA, vA, gold["probe_x"]), gold["probe_Ax"])
    assert bitequal(oracle.transvec(IA, JA, vA, n, gold["probe_y"]), gold["probe_ATy"])
    assert bitequal(oracle.row_norms(IA, vA, 1), gold["rownorm1"])
    assert bitequal(oracle.row_norms(IA, vA, 2), gold["rownorm2"])


def test_c1_phantom(oracle, c1):
    from fortran_airtools_amd import configs as cf
    gold = c1[0]
    assert bitequal(oracle.shepplogan(64), gold["xt"])
    assert bitequal(cf.shepplogan(64), gold["xt"])


@pytest.mark.parametrize("k", [1, 10, 100])
def test_c1_sart(oracle, c1, k):
    gold, IA, JA, vA, n = c1
    assert bitequal(oracle.sirt("sart", IA, JA, vA, n, gold["b"], k, lb=0.0, ub=1.0), gold[f"sart_{k}"])


def test_c1_sart_relax_nobox(oracle, c1):
    gold, IA, JA, vA, n = c1
    assert bitequal(oracle.sirt("sart", IA, JA, vA, n, gold["b"], 20, relax=1.5), gold["sart_norelax_box_none_20"])


@pytest.mark.parametrize("k", [1, 100])
def test_c1_cimmino(oracle, c1, k):
    gold, IA, JA, vA, n = c1
    assert bitequal(oracle.sirt("cimmino", IA, JA, vA, n, gold["b"], k, lb=0.0, ub=1.0), gold[f"cimmino_{k}"])


def test_c1_gd_and_normest(oracle, c1):
    gold, IA, JA, vA, n = c1
    sigma, its = oracle.normest(IA, JA, vA, n, gold["normest_v0"])
    assert abs(sigma - float(gold["normest_sigma"])) <= 1e-14 * sigma
    s = float(gold["normest_sigma"])
    x, _ = oracle.gd(IA, JA, vA, n, gold["b"], 100, s, lb=0.0, ub=1.0)
    assert bitequal(x, gold["gd_100"])
    x, its = oracle.gd(IA, JA, vA, n, gold["b"], 10000, s, tol=1e-3)
    assert bitequal(x, gold["gd_tol_1e-3"])


def test_art_family(oracle):
    g = load_golden("art_small.npz")
    IA, JA, vA, b, n = g["IA"], g["JA"], g["vA"], g["b"], len(g["xt"])
    cases = sorted({k.split("/")[0] for k in g.files if k.startswith("case")})
    for c in cases:
        lb, ub = float(g[f"{c}/lb"]), float(g[f"{c}/ub"])
        X, _ = oracle.art(str(g[f"{c}/method"]), IA, JA, vA, n, b, g[f"{c}/kin"], u=g[f"{c}/u"],
                          lb=lb if np.isfinite(lb) else None, ub=ub if np.isfinite(ub) else None,
                          relax=float(g[f"{c}/relax"]), damp=float(g[f"{c}/damp"]),
                          rowprob=bool(g[f"{c}/rowprob"]))
        assert bitequal(X, g[f"{c}/X"]), c
    X, _ = oracle.art("kaczmarz", IA, JA, vA, n, b, [3], order=g["order"], lb=0.0, ub=1.0)
    assert bitequal(X, g["order_X"])


def twin_case_kwargs(g, c):
    lb, ub = float(g[f"{c}/lb"]), float(g[f"{c}/ub"])
    x0, y0 = g[f"{c}/x0"], g[f"{c}/y0"]
    return dict(u=g[f"{c}/u"], x0=x0 if len(x0) else None, y0=y0 if len(y0) else None,
                slack=int(g[f"{c}/slack"]), lb=lb if np.isfinite(lb) else None, ub=ub if np.isfinite(ub) else None,
                relax=float(g[f"{c}/relax"]), damp=float(g[f"{c}/damp"]), rowprob=bool(g[f"{c}/rowprob"]))


def test_twin_art(oracle):
    """twin_art (AIRMs.f90:371-811) against the compiled reference: with the error-gauge rule the
    iterates are bit-exact and the gauge (NORM2, whose last bit is the compiler runtime's) within
    1e-14; the mutual-step rule normalises with NORM2 and solves a 2x2 system from DOT_PRODUCTs,
    so its iterates carry those last bits (bar 1e-12)."""
    g = load_golden("twin_small.npz")
    IA, JA, vA, b, n = g["IA"], g["JA"], g["vA"], g["b"], len(g["xt"])
    cases = sorted({k.split("/")[0] for k in g.files if k.startswith("case")})
    assert len(cases) == 9
    for c in cases:
        rule = str(g[f"{c}/stoprule"])
        xo, fi, Xs, Ys, EG, q = oracle.twin_art(str(g[f"{c}/method"]), rule, IA, JA, vA, n, b, g[f"{c}/kin"],
                                                **twin_case_kwargs(g, c))
        assert fi == int(g[f"{c}/final_it"]), c
        if str(g[f"{c}/method"]) == "randkaczmarz":
            assert q == len(g[f"{c}/u"]), c
        if rule == "errorgauge":
            assert bitequal(xo, g[f"{c}/xout"]), c
            assert bitequal(Xs, g[f"{c}/Xstore"]) and bitequal(Ys, g[f"{c}/Ystore"]), c
            assert np.max(np.abs(EG - g[f"{c}/EG"])) <= 1e-14 * np.max(np.abs(g[f"{c}/EG"])), c
        else:
            ref = g[f"{c}/xout"]
            assert np.linalg.norm(xo - ref) <= 1e-12 * np.linalg.norm(ref), c
            assert np.max(np.abs(EG - g[f"{c}/EG"])) <= 1e-12 * np.max(np.abs(g[f"{c}/EG"])), c


def test_c2_matrix(oracle):
    from fortran_airtools_amd import configs as cf
    gold = load_golden("c2.npz")
    g = cf.C2
    IA, JA, vA = oracle.fanlineartomo(g.n, g.p, g.nth, g.r, g.dw, g.sd, g.theta())
    assert len(JA) == int(gold["nnz"])
    assert sha(IA) == str(gold["IA_sha"]) and sha(JA) == str(gold["JA_sha"]) and sha(vA) == str(gold["vA_sha"])
    b = cf.add_noise(oracle.matvec(IA, JA, vA, cf.grains(g.n, 200, 2)), 0.005, 2)
    assert sha(b) == str(gold["b_sha"])


def test_c3_matrix_and_slice_inputs(oracle):
    """The batched problem's shared matrix (N=256, 360 x 363 fan) and the golden slices'
    right-hand sides, rebuilt by the oracle, equal the reference's by SHA-256."""
    from fortran_airtools_amd import configs as cf
    gold = load_golden("c3.npz")
    g = cf.C3
    IA, JA, vA = oracle.fanlineartomo(g.n, g.p, g.nth, g.r, g.dw, g.sd, g.theta())
    assert sha(IA) == str(gold["IA_sha"]) and sha(JA) == str(gold["JA_sha"]) and sha(vA) == str(gold["vA_sha"])
    gens = (cf.threephases, cf.bubbles, cf.smooth)
    for j in gold["slices"]:
        xt = gens[j % 3](g.n, seed=3 + j)
        assert sha(xt) == str(gold[f"slice{j}/xt_sha"])
        b = cf.add_noise(oracle.matvec(IA, JA, vA, xt), 0.01, 3 + j)
        assert sha(b) == str(gold[f"slice{j}/b_sha"])


def test_c5_small_group_problems_sart(oracle):
    """The uneven batch's three smallest problems: matrices, right-hand sides and 20 SART
    iterations of the oracle against the reference's."""
    from fortran_airtools_amd import configs as cf
    gold = load_golden("c5.npz")
    probs = cf.c5_problems(64, 5)
    gens = (cf.shepplogan, cf.bubbles, cf.smooth, cf.threephases)
    for k in gold["group_problems"]:
        g = probs[k]
        IA, JA, vA = oracle.fanlineartomo(g.n, g.p, g.nth, g.r, g.dw, g.sd, g.theta())
        assert sha(IA) == str(gold[f"p{k}/IA_sha"]) and sha(vA) == str(gold[f"p{k}/vA_sha"])
        x = cf.shepplogan(g.n) if k % 4 == 0 else gens[k % 4](g.n, seed=5 + k)
        b = cf.add_noise(oracle.matvec(IA, JA, vA, x), 0.01, 5 + k)
        assert sha(b) == str(gold[f"p{k}/b_sha"])
        xo = oracle.sirt("sart", IA, JA, vA, g.ncols, b, 20, lb=0.0, ub=1.0)
        assert bitequal(xo, gold[f"p{k}/sart_20"]), k


def test_c5_symmetric_kaczmarz(oracle):
    """Symmetric Kaczmarz, 20 sweeps on the N=256 fan: the oracle against the reference."""
    from fortran_airtools_amd import configs as cf
    gold = load_golden("c5.npz")
    g = cf.C3
    IA, JA, vA = oracle.fanlineartomo(g.n, g.p, g.nth, g.r, g.dw, g.sd, g.theta())
    b = cf.add_noise(oracle.matvec(IA, JA, vA, cf.threephases(g.n, seed=5)), 0.01, 5)
    assert sha(b) == str(gold["kacz/b_sha"])
    X, _ = oracle.art("symkaczmarz", IA, JA, vA, g.ncols, b, [20], lb=0.0, ub=1.0)
    assert bitequal(X[0], gold["kacz/symkaczmarz_20"])
