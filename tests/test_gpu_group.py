"""Groups of independent SIRT problems (uneven sizes, one matrix each) iterated together: every
problem's iterate equals its own sirt run bit for bit, and the oracle's to 1e-12."""
import numpy as np
import pytest

from conftest import assert_close_1e12

pytestmark = pytest.mark.gpu

# (beam, n, p, nth, deg_step): odd and even p, fan and parallel, different sizes
PROBLEMS = [("fan", 16, 23, 30, 12.0), ("fan", 24, 33, 40, 9.0), ("fan", 32, 45, 60, 6.0),
            ("fan", 20, 28, 36, 10.0), ("par", 18, 25, 24, 7.5)]


def build(oracle=None):
    from fortran_airtools_amd import configs as cf, system_generation as sg
    mats, bs, dats = [], [], []
    for k, (beam, n, p, nth, step) in enumerate(PROBLEMS):
        theta = (0.37 + step * np.arange(nth)) / 180.0 * sg.pi
        d = sg.Data(n=n, p=p, nth=nth, theta=theta)
        A = sg.fanlineartomo(d) if beam == "fan" else sg.paralleltomo(d)
        x = [cf.shepplogan, cf.bubbles, cf.smooth, cf.threephases, cf.shepplogan][k](n)
        mats.append(A)
        bs.append(cf.add_noise(A.matvecmul(x), 0.01, 50 + k))
        dats.append(d)
    return mats, bs, dats


@pytest.mark.parametrize("method", ["sart", "cimmino"])
def test_group_equals_single_runs(method, oracle):
    import torch
    from fortran_airtools_amd import airms
    mats, bs, _ = build()
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    G = airms.SirtGroup(mats, method)
    assert G.autotune() in (0, 1, 2, 3, 4)
    db = [torch.tensor(b, dtype=torch.float64, device="cuda") for b in bs]
    dx = [torch.zeros(A.n, dtype=torch.float64, device="cuda") for A in mats]
    G.sirt_dev([t.data_ptr() for t in db], 15, [t.data_ptr() for t in dx], o)
    torch.cuda.synchronize()
    for k, A in enumerate(mats):
        single = torch.zeros(A.n, dtype=torch.float64, device="cuda")
        airms.sirt_dev(method, A, db[k].data_ptr(), 15, single.data_ptr(), o)
        torch.cuda.synchronize()
        assert dx[k].cpu().numpy().tobytes() == single.cpu().numpy().tobytes(), k
    # the oracle on the first problem
    IA, JA, vA = mats[0].components()
    xo = oracle.sirt(method, IA, JA, vA, mats[0].n, bs[0], 15, lb=0.0, ub=1.0)
    assert_close_1e12(dx[0].cpu().numpy(), xo)


def test_group_continues_from_start_vectors():
    """A second call continues from the current iterates: 5 + 10 iterations equal 15."""
    import torch
    from fortran_airtools_amd import airms
    mats, bs, _ = build()
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    G = airms.SirtGroup(mats, "sart")
    db = [torch.tensor(b, dtype=torch.float64, device="cuda") for b in bs]
    x1 = [torch.zeros(A.n, dtype=torch.float64, device="cuda") for A in mats]
    x2 = [torch.zeros(A.n, dtype=torch.float64, device="cuda") for A in mats]
    G.sirt_dev([t.data_ptr() for t in db], 15, [t.data_ptr() for t in x1], o)
    G.sirt_dev([t.data_ptr() for t in db], 5, [t.data_ptr() for t in x2], o)
    G.sirt_dev([t.data_ptr() for t in db], 10, [t.data_ptr() for t in x2], o)
    torch.cuda.synchronize()
    for a, b in zip(x1, x2):
        assert a.cpu().numpy().tobytes() == b.cpu().numpy().tobytes()


def test_spmv_autotune_keeps_results():
    """Re-tiling the staged layouts changes no result: rows keep their accumulation order."""
    import ctypes as C
    from fortran_airtools_amd import _lib, airms, configs as cf, system_generation as sg
    g = cf.C1
    A = sg.fanlineartomo(sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta()))
    x = np.random.default_rng(1).random(A.n)
    y = np.random.default_rng(2).random(A.m)
    b = A.matvecmul(cf.shepplogan(g.n))
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    before = (A.matvecmul(x), A.transvecmul(y), airms.sirt("sart", A, b, [5], options=o))
    choice = (C.c_int32 * 6)()
    _lib.check(_lib.lib().fair_autotune_spmv(A.handle, choice))
    assert list(choice)[0] >= 0
    after = (A.matvecmul(x), A.transvecmul(y), airms.sirt("sart", A, b, [5], options=o))
    for u, v in zip(before, after):
        assert u.tobytes() == v.tobytes()


def test_store_and_read_device_matrix(tmp_path, oracle):
    """A device-built matrix stored with store_matrix and read back with read_csr_matrix gives
    the same products (rows as the caller sees them: sorted after SART, as in the reference)."""
    from fortran_airtools_amd import airms, configs as cf, sparse_matrices as sm, system_generation as sg
    g = cf.C1
    A = sg.fanlineartomo(sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta()))
    airms.sirt("sart", A, A.matvecmul(cf.shepplogan(g.n)), [1])     # sorts the caller's rows
    path = str(tmp_path / "A.bin")
    sm.store_matrix(A, path)
    B = sm.read_csr_matrix(path)
    IA, JA, vA = A.components()
    IB, JB, vB = B.components()
    assert IA.tobytes() == IB.tobytes() and JA.tobytes() == JB.tobytes() and vA.tobytes() == vB.tobytes()
    # The device keeps its own entry order inside a row (and a staged layout) for A; B holds the
    # sorted rows of the file: the same sums in another order, equal to the parity bar.
    x = np.random.default_rng(5).random(A.n)
    want = oracle.matvec(IB, JB, vB, x)
    assert_close_1e12(A.matvecmul(x), want)
    assert_close_1e12(B.matvecmul(x), want)
