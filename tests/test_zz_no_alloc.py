"""fair_sirt_dev_noalloc (sirt with its two workspaces in the library's static device array,
no device allocation) against fair_sirt_dev (workspaces allocated on the matrix): the iterates
are the same bit for bit -- the two entry points run the same kernels -- on two calls in a row,
the second on a smaller problem (the array keeps the first call's contents), and equal the
oracle's sirt (AIRMs.f90:997-1127) to the 1e-12 bar.  An unprepared matrix and a problem beyond
the array are refused with FAIR_ERR_ARG before anything runs."""
import ctypes as C

import numpy as np
import pytest

from conftest import assert_close_1e12

pytestmark = pytest.mark.gpu

METHODS = {"cimmino": 0, "sart": 1}


def fan(n, p, nth, step):
    from fortran_airtools_amd import system_generation as sg
    theta = np.deg2rad(np.arange(nth) * step)
    return sg.fanlineartomo(sg.Data(n=n, p=p, nth=nth, theta=theta))


def both(A, method, b, iters, o):
    """(fair_sirt_dev's iterate, fair_sirt_dev_noalloc's iterate) from x0 = 0."""
    import torch
    from fortran_airtools_amd import _lib
    L = _lib.lib()
    oc = o._c()
    db = torch.from_numpy(b).cuda()
    x_alloc = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    x_static = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    _lib.check(L.fair_sirt_prepare(A.handle, METHODS[method]))
    torch.cuda.synchronize()
    _lib.check(L.fair_sirt_dev(A.handle, METHODS[method], C.c_void_p(db.data_ptr()), iters,
                               C.c_void_p(x_alloc.data_ptr()), C.byref(oc)))
    torch.cuda.synchronize()
    _lib.check(L.fair_sirt_dev_noalloc(A.handle, METHODS[method], C.c_void_p(db.data_ptr()), iters,
                                       C.c_void_p(x_static.data_ptr()), C.byref(oc)))
    torch.cuda.synchronize()
    return x_alloc.cpu().numpy(), x_static.cpu().numpy()


def test_noalloc_equals_sirt_dev_on_two_calls(oracle):
    from fortran_airtools_amd import airms, configs as cf
    cases = [(fan(64, 91, 180, 2.0), "sart", 20, dict(lb=0.0, ub=1.0, relax=1.0)),
             (fan(24, 33, 40, 9.0), "cimmino", 15, dict(lb=None, ub=1.0, relax=1.3))]
    for k, (A, method, iters, kw) in enumerate(cases):
        print("call", k, method, A.m, A.n, flush=True)
        b = cf.add_noise(A.matvecmul(cf.shepplogan(int(round(np.sqrt(A.n))))), 0.01, 30 + k)
        o = airms.Opts()
        o.relaxpar = kw["relax"]
        if kw["lb"] is not None:
            o.set_lbound(kw["lb"])
        if kw["ub"] is not None:
            o.set_ubound(kw["ub"])
        x_alloc, x_static = both(A, method, b, iters, o)
        assert x_alloc.tobytes() == x_static.tobytes(), k
        IA, JA, vA = A.components()
        xo = oracle.sirt(method, IA, JA, vA, A.n, b, iters, lb=kw["lb"], ub=kw["ub"], relax=kw["relax"])
        assert_close_1e12(x_static, xo)


def test_noalloc_refuses_unprepared_and_oversized_matrices():
    import torch
    from fortran_airtools_amd import _lib, airms
    L = _lib.lib()
    oc = airms.Opts()._c()
    A = fan(16, 21, 36, 10.0)
    b = torch.zeros(A.m, dtype=torch.float64, device="cuda")
    x = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    st = L.fair_sirt_dev_noalloc(A.handle, 1, C.c_void_p(b.data_ptr()), 3, C.c_void_p(x.data_ptr()), C.byref(oc))
    assert st == _lib.FAIR_ERR_ARG
    assert b"fair_sirt_prepare" in L.fair_last_error()
    big = fan(2048, 5, 2, 90.0)        # n = 2048^2 doubles alone fill the array
    st = L.fair_sirt_dev_noalloc(big.handle, 1, C.c_void_p(b.data_ptr()), 3, C.c_void_p(x.data_ptr()), C.byref(oc))
    assert st == _lib.FAIR_ERR_ARG
    assert b"exceeds the static SIRT scratch" in L.fair_last_error()
