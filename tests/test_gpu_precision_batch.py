"""Precision variant (FP32 values / iterate / accumulation, FP64 weights) and batched SIRT.

Bar: the FP32 iterates agree with the plain FP64 method to 1e-5 relative (norm-wise), the
batched run agrees with per-slice FP32 runs to 1e-6, and graph replays reproduce the eager
batched result bit for bit.
"""
import numpy as np
import pytest

from conftest import load_golden, relerr

pytestmark = pytest.mark.gpu


def c1():
    from fortran_airtools_amd import configs as cf, system_generation as sg
    g = cf.C1
    return sg.fanlineartomo(sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta()))


@pytest.mark.parametrize("method", ["sart", "cimmino"])
def test_fp32_sirt_within_1e5_of_fp64(method):
    from fortran_airtools_amd import airms
    gold = load_golden("c1.npz")
    A = c1()
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    x32 = airms.sirt_f32(method, A, gold["b"], [100], options=o)
    assert x32.dtype == np.float32
    assert relerr(x32, gold[f"{method}_100"]) <= 1e-5


def batch_inputs(A, k):
    from fortran_airtools_amd import configs as cf
    gen = [cf.shepplogan, cf.bubbles, cf.smooth, cf.threephases, cf.grains]
    N = int(round(np.sqrt(A.n)))
    X = [gen[j % len(gen)](N) for j in range(k)]
    B = np.stack([cf.add_noise(A.matvecmul(x), 0.01, 100 + j) for j, x in enumerate(X)], 1)
    return B


def test_batch_matches_single_and_fp64():
    from fortran_airtools_amd import airms
    A = c1()
    k = 20
    B = batch_inputs(A, k)
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    X = airms.sirt_batch("sart", A, B, [30], options=o)
    assert X.shape == (A.n, k)
    for j in (0, 7, 19):
        xs = airms.sirt_f32("sart", A, B[:, j], [30], options=o)
        assert relerr(X[:, j], xs) <= 1e-6
        x64 = airms.sirt("sart", A, B[:, j], [30], options=o)
        assert relerr(X[:, j], x64) <= 1e-5


def test_batch_graph_replays_match_eager():
    from fortran_airtools_amd import airms
    A = c1()
    k = 33                      # not a multiple of the 16-wide chunk
    B = batch_inputs(A, k)
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    eager = airms.sirt_batch("cimmino", A, B, [12], options=o)
    for rep in range(3):
        # eager work between replays
        A.matvecmul(np.ones(A.n))
        g = airms.sirt_batch("cimmino", A, B, [12], options=o, use_graph=True)
        assert g.tobytes() == eager.tobytes(), rep


def test_batch_graph_survives_workspace_and_matrix_changes():
    """Graph replays of fair_sirt_batch_dev after changes to what a captured run refers to: a
    larger k (the workspaces are reallocated), the earlier k again, and the matrix scaled by
    multiply_by_scalar (FP32 values and weights rebuilt).  Every graph result equals the eager
    run of the same call bit for bit."""
    import torch
    from fortran_airtools_amd import airms
    A = c1()
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    B = {k: torch.from_numpy(batch_inputs(A, k).astype(np.float32)).cuda() for k in (8, 16)}
    # one output panel per k: a graph captured for (B[k], X[k]) is found again by later calls
    X = {k: torch.zeros((A.n, k), dtype=torch.float32, device="cuda") for k in (8, 16)}

    def run(k, graph):
        X[k].zero_()
        torch.cuda.synchronize()
        airms.sirt_batch_dev("sart", A, B[k].data_ptr(), k, 10, X[k].data_ptr(), options=o, use_graph=graph)
        torch.cuda.synchronize()
        return X[k].cpu().numpy()

    for step, k in enumerate((8, 16, 8)):
        g = run(k, True)
        assert g.tobytes() == run(k, False).tobytes(), (step, k)
    first = run(8, True)
    A.multiply_by_scalar(0.5)
    g = run(8, True)
    assert g.tobytes() == run(8, False).tobytes()
    assert g.tobytes() != first.tobytes()


@pytest.mark.parametrize("geom", ["fan_c1", "parallel_n96_pi"])
@pytest.mark.parametrize("tile_code", [100, 101, 102, 103, 104, 105, 110, 111, 112, 113, 114, 115,
                                       120, 121, 123, 125, 130, 133, 135])
def test_tiled_ray_product_equals_row_kernel(monkeypatch, geom, tile_code):
    """The LDS-staged tiled kernel for A*X (rays of a sinogram tile of 8 x 8 to 4 x 4 angles x
    detectors, their pixels' union staged in an order every ray's stored order follows; tiles
    without such an order are split; codes 12x / 13x stage slices of 64 right-hand sides) gives the
    same bits as the wavefront-per-row kernel.  The fan
    matrix has tiles that need splitting; the parallel-beam one has columns outside the image, for
    which the layout does not apply and the row kernel runs.  k = 32 and 256."""
    import torch
    from fortran_airtools_amd import airms
    A = c1() if geom == "fan_c1" else parallel_pi()
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    for k in (32, 256):
        B = torch.from_numpy(batch_inputs(A, k).astype(np.float32)).cuda()
        out = {}
        for codes in ("3,6", f"{tile_code},6"):
            monkeypatch.setenv("FAIR_SPMM_ORDERS", codes)
            X = torch.zeros((A.n, k), dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            airms.sirt_batch_dev("sart", A, B.data_ptr(), k, 12, X.data_ptr(), options=o)
            torch.cuda.synchronize()
            out[codes] = X.cpu().numpy()
        assert np.abs(out["3,6"]).max() > 0, k
        assert out["3,6"].tobytes() == out[f"{tile_code},6"].tobytes(), k


def test_batch_graph_after_row_order_changes(monkeypatch):
    """A changed row order or tiled layout replaces the matrix's cached copy, which captured runs
    refer to: graph runs over orders 3,6 -> 2,100 (tiled A^T) -> 2002,10 (four rows per wavefront)
    -> 32002,40010 (with residency caps) -> 3,6 again each equal the eager run bit for bit."""
    import torch
    from fortran_airtools_amd import airms
    A = c1()
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    k = 256
    B = torch.from_numpy(batch_inputs(A, k).astype(np.float32)).cuda()
    X = torch.zeros((A.n, k), dtype=torch.float32, device="cuda")

    def run(graph):
        X.zero_()
        torch.cuda.synchronize()
        airms.sirt_batch_dev("sart", A, B.data_ptr(), k, 6, X.data_ptr(), options=o, use_graph=graph)
        torch.cuda.synchronize()
        return X.cpu().numpy()

    monkeypatch.setenv("FAIR_SPMM_ORDERS", "3,6")
    eager = run(False)
    for codes in ("3,6", "2,100", "2002,10", "32002,40010", "3,6"):
        monkeypatch.setenv("FAIR_SPMM_ORDERS", codes)
        assert run(True).tobytes() == eager.tobytes(), codes
        assert run(True).tobytes() == eager.tobytes(), codes


@pytest.mark.parametrize("tile_code", [100, 101, 102, 103, 104, 105, 106, 107, 108,
                                       110, 111, 112, 113, 114, 115, 116, 117, 118,
                                       120, 125, 127, 128, 130, 135, 137])
def test_tiled_transpose_product_equals_row_kernel(monkeypatch, tile_code):
    """The LDS-staged tiled kernel of A^T*R (pixels of an image tile, their rays' union staged in
    double-buffered chunks of 64 (codes 100-108) or 32 (110-118) rows; 16-pixel tiles one pixel per
    wavefront, 8 x 4 and 4 x 8 tiles (codes x05, x06) two, 8 x 8 and 16 x 4 (x07, x08) four; codes
    12x and 13x stage slices of 64 right-hand sides in chunks of 128 and 64 rows, one pixel per
    16-lane group; k = 32 leaves half of each group idle) gives the same bits as the
    wavefront-per-row kernel: every pixel's sum runs over its entries in the stored order with the
    same operations.  k = 32 and k = 256 right-hand sides."""
    import torch
    from fortran_airtools_amd import airms
    A = c1()
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    for k in (32, 256):
        B = torch.from_numpy(batch_inputs(A, k).astype(np.float32)).cuda()
        out = {}
        for codes in ("3,6", f"3,{tile_code}"):
            monkeypatch.setenv("FAIR_SPMM_ORDERS", codes)
            X = torch.zeros((A.n, k), dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            airms.sirt_batch_dev("sart", A, B.data_ptr(), k, 12, X.data_ptr(), options=o)
            torch.cuda.synchronize()
            out[codes] = X.cpu().numpy()
        assert out["3,6"].tobytes() == out[f"3,{tile_code}"].tobytes(), k


def parallel_pi():
    """Parallel beam, N = 96, 90 angles up to exactly pi: the tracer emits 77 column numbers above
    N^2 at the last angle (a reference quirk kept in the matrix), which the products read as zero."""
    from fortran_airtools_amd import system_generation as sg
    return sg.paralleltomo(sg.Data(n=96, p=137, nth=90, theta=np.deg2rad(np.linspace(0.0, 180.0, 90))))


@pytest.mark.parametrize("geom", ["fan_c1", "parallel_n96_pi"])
@pytest.mark.parametrize("grouping", [1, 2])
def test_grouped_rows_equal_one_row_per_wavefront(monkeypatch, geom, grouping):
    """Both products with two rows per wavefront in passes of 128 right-hand sides (grouping 1) or
    four rows in passes of 64 (grouping 2), also with a cap on the workgroups resident per compute
    unit, give the same bits as one row per wavefront over all of them: each output sums the same
    entries in the same order.  k = 96 (grouping 1 does not apply
    and runs one row per wavefront), 132 (a last pass of 4) and 256; row lengths differ inside a
    wavefront, and the parallel-beam matrix has columns outside the image."""
    import torch
    from fortran_airtools_amd import airms
    A = c1() if geom == "fan_c1" else parallel_pi()
    if geom == "parallel_n96_pi":
        assert int((np.asarray(A.JA) >= A.n).sum()) > 0
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    for k in (96, 132, 256):
        B = torch.from_numpy(batch_inputs(A, k).astype(np.float32)).cuda()
        out = {}
        group_codes = f"{1000 * grouping + 3},{1000 * grouping + 6}"
        capped = f"{30000 + 1000 * grouping + 3},{50006}"   # at most 3 / 5 workgroups per CU
        for codes in ("3,6", group_codes, capped):
            monkeypatch.setenv("FAIR_SPMM_ORDERS", codes)
            X = torch.zeros((A.n, k), dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            airms.sirt_batch_dev("sart", A, B.data_ptr(), k, 12, X.data_ptr(), options=o)
            torch.cuda.synchronize()
            out[codes] = X.cpu().numpy()
        assert np.isfinite(out["3,6"]).all() and np.abs(out["3,6"]).max() > 0, k
        assert out["3,6"].tobytes() == out[group_codes].tobytes(), k
        assert out["3,6"].tobytes() == out[capped].tobytes(), k
