"""Benchmark: SART iterations per second and CSR SpMV HBM bandwidth, N=512 fan beam, FP64.

Workload (one step = one SART iteration, the reference's `sirt(...,'sart',...)` inner loop,
AIRMs.f90:1089-1110): 512x512 pixels, fan beam with 720 angles (0:0.5:359.5 deg) x 725 rays
(r=2, dw=3.1, sd=4 in units of N), system matrix built on the device (CSR FP64 values, int32
indices), seeded Voronoi "grains" phantom, 0.5 % Gaussian noise, box [0,1], relax 1.
Everything is resident in HBM before the timed region.

Prints one JSON line (rank 0, one GPU).  Extra objects: `roofline` for the SpMV kernel
(algorithmic bytes 12*nnz + 4*(m+1) + 8*n + 8*m per launch over its HIP-event-timed average
launch duration) and `cpu_baseline` (the reference's own OpenMP SART timed on this host).

usage: python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SART iter/s + SpMV HBM GB/s (N=512 fan); batched FP32 vs FP64 at equal error"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measured_traffic() -> int | None:
    """HBM bytes per SpMV launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    (profiles/spmv_n512_fan_fp64_hbm_traffic.md, line "traffic_bytes_per_launch: <int>")."""
    path = os.path.join(ROOT, "profiles", "spmv_n512_fan_fp64_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def matfree_counts() -> dict:
    """Per-segment flop and VALU-operation counts of the matrix-free kernels from the committed
    counter summary (profiles/matrix_free_fan_n2048_fp32_valu_counters.md, "name: value" lines)."""
    path = os.path.join(ROOT, "profiles", "matrix_free_fan_n2048_fp32_valu_counters.md")
    out = {}
    try:
        with open(path) as f:
            for line in f:
                k, sep, v = line.partition(":")
                if sep and k.strip().endswith("_per_segment"):
                    out[k.strip()] = float(v)
    except (OSError, ValueError):
        pass
    return out


def run_cpu_baseline(geom=None, legs=("sart@16",), time_build=False, iters: int = 3) -> dict | None:
    """The reference's own methods (oracle/_ref, OpenMP) on a fan-beam problem (default N=512,
    720 x 725 rays; `geom` a configs.Geometry), one leg per "method@threads" (threads capped at
    the 16 host cores a one-GPU box gives a job, and at the affinity count), each timed over a
    bounded number of iterations; with time_build also the reference's serial fanlineartomo."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cap = max(1, min(avail, 16))
    fixed = []
    for leg in legs:
        meth, _, thr = leg.partition("@")
        fixed.append(f"{meth}@{min(int(thr or 1), cap)}")
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--iters", str(iters), "--legs", ",".join(fixed)]
    if time_build:
        cmd.append("--time-build")
    if geom is not None:
        cmd += ["--n", str(geom.n), "--p", str(geom.p), "--nth", str(geom.nth), "--step", str(geom.deg_step),
                "--r", str(geom.r), "--dw", str(geom.dw), "--sd", str(geom.sd)]
    log("cpu baseline:", " ".join(cmd))
    import resource

    def big_stack():
        # the reference keeps m- and n-sized automatic arrays on the stack (AIRMs.f90:1009-1010)
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    env = dict(os.environ, OMP_STACKSIZE="256M")
    try:
        # its progress lines pass through on stderr; the result is the last stdout line that holds
        # a JSON object (the reference's own output, flushed at exit, may follow it)
        r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, text=True, timeout=600,
                           preexec_fn=big_stack, env=env)
    except subprocess.TimeoutExpired:
        log("cpu baseline timed out")
        return None
    if r.returncode != 0:
        log(f"cpu baseline failed (status {r.returncode})")
        return None
    for line in reversed(r.stdout.strip().splitlines()):
        if line.startswith("{"):
            return json.loads(line)
    log("cpu baseline printed no result")
    return None


def leg(d: dict | None, i: int) -> dict | None:
    """The i-th asked leg of a run_cpu_baseline result (None when absent)."""
    if not d:
        return None
    vals = list(d["legs"].values())
    return vals[i] if i < len(vals) else None


def cpu_baseline(iters: int, method: str = "sart", geom=None) -> dict | None:
    """One leg (the method on up to 16 host threads) in the bench line's cpu_baseline form."""
    d = run_cpu_baseline(geom, (f"{method}@16",), iters=iters)
    lg = leg(d, 0)
    if lg is None:
        return None
    return {"value": round(lg["iter_per_s"], 4), "unit": "iter/s", "cores": lg["cores"], "kind": d["kind"],
            "sample": lg["sample"] + f"; {d['cpu_model']}",
            "gbs": None if lg["gbs"] is None else round(lg["gbs"], 2),
            "nproc": d["nproc"], "affinity": d["affinity"]}


def eq_precision(torch, _lib, L, A, b, oc, stream, iters: int = 200) -> dict:
    """FP32 (values, iterate, accumulation; FP64 weights) against the plain FP64 SART, both from
    x0 = 0 for `iters` iterations on HBM-resident inputs: device time of each (HIP events on the
    library's stream) and the relative error of the FP32 iterate to the FP64 one."""
    dev = torch.device("cuda:0")
    b64 = torch.from_numpy(b).to(dev)
    b32 = b64.float()
    x64 = torch.zeros(A.n, dtype=torch.float64, device=dev)
    x32 = torch.zeros(A.n, dtype=torch.float32, device=dev)
    _lib.check(L.fair_sirt_f32_dev(A.handle, 1, C.c_void_p(b32.data_ptr()), 1, C.c_void_p(x32.data_ptr()),
                                   C.byref(oc)))          # FP32 values and workspaces, untimed
    x32.zero_()
    torch.cuda.synchronize()
    e0, e1 = _events(torch, stream)
    e0.record(stream)
    _lib.check(L.fair_sirt_dev(A.handle, 1, C.c_void_p(b64.data_ptr()), iters, C.c_void_p(x64.data_ptr()),
                               C.byref(oc)))
    e1.record(stream)
    torch.cuda.synchronize()
    t64 = e0.elapsed_time(e1) / 1e3
    e0.record(stream)
    _lib.check(L.fair_sirt_f32_dev(A.handle, 1, C.c_void_p(b32.data_ptr()), iters, C.c_void_p(x32.data_ptr()),
                                   C.byref(oc)))
    e1.record(stream)
    torch.cuda.synchronize()
    t32 = e0.elapsed_time(e1) / 1e3
    err = float(torch.linalg.norm(x32.double() - x64) / torch.linalg.norm(x64))
    log(f"C2 SART {iters} it: FP64 {t64 * 1e3:.2f} ms, FP32 {t32 * 1e3:.2f} ms, rel. error {err:.2e}")
    return {"workload": f"N=512 fan 720 x 725, SART {iters} iterations, box [0,1]", "iterations": iters,
            "fp64_s": round(t64, 6), "fp32_s": round(t32, 6), "fp32_relerr_vs_fp64": err,
            "within_tolerance": err <= 1e-5, "speedup": round(t64 / t32, 3)}


def eq_batch(torch, _lib, L, airms, cf, sg, stream, iters: int = 300, k: int = 256) -> dict:
    """256 slices of 256^2 sharing one fan-beam matrix (360 x 363 rays): one batched FP32 SART run
    (SpMM, tuned row orders, HIP graph) against the plain FP64 method slice by slice; per-slice
    relative error of the batched iterate to the FP64 one."""
    g = cf.C3
    A = sg.fanlineartomo(sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta()))
    gens = [cf.threephases, cf.bubbles, cf.smooth]
    dev = torch.device("cuda:0")
    B = np.stack([cf.add_noise(A.matvecmul(gens[j % 3](g.n, seed=3 + j)), 0.01, 3 + j) for j in range(k)], 1)
    oc = airms.Opts().set_lbound(0.0).set_ubound(1.0)._c()
    dB = torch.from_numpy(B.astype(np.float32)).to(dev)
    dX = torch.zeros((A.n, k), dtype=torch.float32, device=dev)
    _lib.check(L.fair_sirt_prepare(A.handle, 1))
    order = (C.c_int32 * 2)()
    _lib.check(L.fair_autotune_batch(A.handle, k, order))
    torch.cuda.synchronize()
    for _ in range(2):      # capture (first call) and one replay, untimed
        dX.zero_()
        torch.cuda.synchronize()
        _lib.check(L.fair_sirt_batch_dev(A.handle, 1, C.c_void_p(dB.data_ptr()), k, iters,
                                         C.c_void_p(dX.data_ptr()), C.byref(oc), 1))
    dX.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check(L.fair_sirt_batch_dev(A.handle, 1, C.c_void_p(dB.data_ptr()), k, iters,
                                     C.c_void_p(dX.data_ptr()), C.byref(oc), 1))
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    dB64 = torch.from_numpy(B).to(dev)
    Xref = torch.zeros((A.n, k), dtype=torch.float64, device=dev)
    col = torch.zeros(A.n, dtype=torch.float64, device=dev)
    bcol = torch.zeros(A.m, dtype=torch.float64, device=dev)
    t64 = 0.0
    e0, e1 = _events(torch, stream)
    for j in range(k):
        bcol.copy_(dB64[:, j])
        col.zero_()
        torch.cuda.synchronize()
        e0.record(stream)
        _lib.check(L.fair_sirt_dev(A.handle, 1, C.c_void_p(bcol.data_ptr()), iters, C.c_void_p(col.data_ptr()),
                                   C.byref(oc)))
        e1.record(stream)
        torch.cuda.synchronize()
        t64 += e0.elapsed_time(e1) / 1e3
        Xref[:, j] = col
    err = np.empty(k)
    _lib.check(L.fair_batch_relerr_dev(C.c_void_p(dX.data_ptr()), C.c_void_p(Xref.data_ptr()), A.n, k,
                                       err.ctypes.data_as(C.c_void_p)))
    nnz = A.nnz
    log(f"C3 batched FP32 SART {k} slices x {iters} it (graph): {t_batch:.3f} s; FP64 slice by slice "
        f"{t64:.3f} s; rel. error max {err.max():.2e}")
    return {"workload": f"{k} slices of 256x256, fan 360 x 363, SART as SpMM, {iters} iterations, HIP graph",
            "slices": k, "iterations": iters, "nnz": nnz, "batched_fp32_s": round(t_batch, 4),
            "fp64_slice_by_slice_s": round(t64, 4), "speedup": round(t64 / t_batch, 3),
            "ms_per_batched_iteration": round(t_batch / iters * 1e3, 4),
            "relerr_vs_fp64": {"max": float(err.max()), "median": float(np.median(err))},
            "within_tolerance": bool((err <= 1e-5).all()),
            "achieved_tflops_fp32": round(4 * nnz * k * iters / t_batch / 1e12, 3),
            "row_orders": [order[0], order[1]]}


def eq_matfree_n512(torch, airms, cf, sg, A, iters: int = 100) -> dict:
    """Gradient descent on the N=512 problem (720 x 725 rays, grains phantom, box [0,1]): the
    plain FP64 CSR method against the matrix-free FP32 projector from the same sigma, the
    relative error of the matrix-free iterate to the FP64 one."""
    from fortran_airtools_amd.matrix_free import FanProjector
    g = cf.C2
    dev = torch.device("cuda:0")
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    b = A.matvecmul(cf.grains(g.n, 200, 2))
    sigma = airms.normest(A, v0=np.random.default_rng(2).random(A.n))
    airms.gradient_descent(A, b, 1, tolerance=0.0, options=o, sigma=sigma)      # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x64 = airms.gradient_descent(A, b, iters, tolerance=0.0, options=o, sigma=sigma)
    t64 = time.perf_counter() - t0
    P = FanProjector(sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta()))
    db = torch.from_numpy(b.astype(np.float32)).to(dev)
    x32 = torch.zeros(A.n, dtype=torch.float32, device=dev)
    P.gradient_descent_dev(db.data_ptr(), x32.data_ptr(), max_iter=1, tolerance=0.0, sigma=sigma, options=o)
    x32.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P.gradient_descent_dev(db.data_ptr(), x32.data_ptr(), max_iter=iters, tolerance=0.0, sigma=sigma, options=o)
    torch.cuda.synchronize()
    t32 = time.perf_counter() - t0
    err = float(np.linalg.norm(x32.cpu().numpy().astype(np.float64) - x64) / np.linalg.norm(x64))
    P.close()
    log(f"N=512 GD {iters} it: FP64 CSR {t64:.3f} s, matrix-free FP32 {t32:.3f} s, rel. error {err:.2e}")
    return {"workload": f"N=512 fan 720 x 725, gradient descent {iters} iterations, box [0,1]",
            "iterations": iters, "fp64_csr_s": round(t64, 4), "matrix_free_fp32_s": round(t32, 4),
            "relerr_vs_fp64": err, "within_tolerance": err <= 1e-5}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-equal-error", action="store_true",
                    help="skip the time-to-equal-error runs (FP32 at N=512, the 256-slice batch, matrix-free GD)")
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--mode", default="sart",
                    choices=["sart", "precision", "batch", "matfree", "kaczmarz", "uneven", "phantoms"],
                    help="sart: the headline N=512 FP64 SART line (default); precision: FP32 vs FP64 "
                         "time to equal error at N=512; batch: 256 slices of 256^2 as one batched FP32 run; "
                         "matfree: matrix-free FP32 gradient descent at 2048^2 and at N=512 against the "
                         "FP64 CSR method; kaczmarz: randomized (row-norm sampling) and symmetric "
                         "Kaczmarz, 20 sweeps at N=256, cyclic Kaczmarz and twin_art, against the oracle and the "
                         "reference on the host; "
                         "uneven: 64 SART problems from N=128 to N=1024 as one group against one run each; "
                         "phantoms: every phantom of the gallery and add_white_noise at N=2048 and N=256, "
                         "the reference's own functions timed on the host at N=256")
    ap.add_argument("--uneven-count", type=int, default=64)
    ap.add_argument("--uneven-sizes", default="", help="comma-separated image sizes for --mode uneven "
                    "(default: the C5 set, 128..1024)")
    ap.add_argument("--batch-iters", type=int, default=300)
    ap.add_argument("--mf-iters", type=int, default=100)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 1 GiB fill and a 1 GiB sum (reference kernels for HBM counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: the library drives a single device and has no "
              f"multi-GPU path (asked for --gpus {a.gpus}).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from fortran_airtools_amd import _lib, airms, configs as cf, system_generation as sg

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    L = _lib.lib()
    stream = torch.cuda.current_stream()
    _lib.check(L.fair_set_stream(C.c_void_p(stream.cuda_stream)))
    if a.mode == "precision":
        return bench_precision(a, np, torch, _lib, airms, cf, sg, L, stream)
    if a.mode == "batch":
        return bench_batch(a, np, torch, _lib, airms, cf, sg, L, stream)
    if a.mode == "matfree":
        return bench_matfree(a, np, torch, _lib, airms, cf, sg, L, stream)
    if a.mode == "kaczmarz":
        return bench_kaczmarz(a, np, torch, _lib, airms, cf, sg, L, stream)
    if a.mode == "uneven":
        return bench_uneven(a, np, torch, _lib, airms, cf, sg, L, stream)
    if a.mode == "phantoms":
        return bench_phantoms(a, np, torch)

    g = cf.C2
    dat = sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta())
    # the first build of a process also loads the library's code object onto the device: timed
    # apart from a second, warm build (the reference's build is timed in a warm process too)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A = sg.fanlineartomo(dat)
    torch.cuda.synchronize()
    t_build_first = time.perf_counter() - t0
    del A
    t0 = time.perf_counter()
    A = sg.fanlineartomo(dat)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    m, n, nnz = A.m, A.n, A.nnz
    log(f"matrix {m} x {n}, nnz {nnz}, device build {t_build:.4f} s (first call of the process "
        f"{t_build_first:.3f} s)")

    xt = cf.grains(g.n, 200, 2)
    b = cf.add_noise(A.matvecmul(xt), 0.005, 2)
    opts = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    oc = opts._c()
    _lib.check(L.fair_sirt_prepare(A.handle, 1))

    d_b = torch.from_numpy(b).to(dev)
    d_x = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def sart(k):
        _lib.check(L.fair_sirt_dev(A.handle, 1, C.c_void_p(d_b.data_ptr()), int(k),
                                   C.c_void_p(d_x.data_ptr()), C.byref(oc)))

    # the SpMV kernel alone (MMRS over the non-empty rows), HIP events on the launch stream
    xs = torch.rand(n, dtype=torch.float64, device=dev)
    ys = torch.zeros(m, dtype=torch.float64, device=dev)

    def spmv():
        _lib.check(L.fair_spmv_rows_dev(A.handle, C.c_void_p(xs.data_ptr()), C.c_void_p(ys.data_ptr())))

    # bring the device to its steady clocks before anything is timed: about 0.2 s of SpMV launches
    # (an idle GPU runs the first milliseconds of work slower)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.2:
        for _ in range(50):
            spmv()
        torch.cuda.synchronize()

    sart(a.warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    sart(a.steps)
    e1.record(stream)
    torch.cuda.synchronize()
    t_sart = e0.elapsed_time(e1) / 1e3
    it_s = a.steps / t_sart
    bytes_it = 24 * nnz + 4 * (m + n + 2) + 32 * m + 32 * n
    log(f"SART {a.steps} iterations: {t_sart * 1e3:.2f} ms, {it_s:.1f} it/s, "
        f"{bytes_it / (t_sart / a.steps) / 1e9:.1f} GB/s per iteration")

    # the same iterations through the host-array entry point (b and x0 cross PCIe in, x out)
    t0 = time.perf_counter()
    airms.sirt("sart", A, b, [a.steps], x0=np.zeros(n), options=opts)
    t_host = time.perf_counter() - t0
    log(f"SART {a.steps} iterations from host arrays (PCIe-inclusive): {t_host * 1e3:.2f} ms")

    for _ in range(a.warmup):
        spmv()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(a.steps):
        spmv()
    e1.record(stream)
    torch.cuda.synchronize()
    t_spmv = e0.elapsed_time(e1) / 1e3 / a.steps
    bytes_spmv = 12 * nnz + 4 * (m + 1) + 8 * n + 8 * m
    gbs = bytes_spmv / t_spmv / 1e9
    log(f"SpMV: {t_spmv * 1e6:.1f} us/launch, {gbs:.1f} GB/s ({gbs / HBM_PEAK_GBS:.1%} of peak)")

    if a.calibrate:
        # reference kernels of known HBM traffic for the FETCH_SIZE / WRITE_SIZE counters:
        # a 1 GiB fill (writes 1 GiB) and a 1 GiB sum (reads 1 GiB; larger than the MALL)
        big = torch.full((128 << 20,), 1.0, dtype=torch.float64, device=dev)
        for _ in range(2):
            float(big.sum())
        torch.cuda.synchronize()
        del big

    # time to equal error (<= 1e-5 relative to the plain FP64 iterate): FP32 at N=512, the batched
    # FP32 run of 256 slices, matrix-free FP32 gradient descent at N=512
    eq = None
    if not a.no_equal_error:
        eq = {"tolerance": 1e-5, "reference": "plain FP64 device method at the same iteration count"}
        eq["c2_fp32_vs_fp64"] = eq_precision(torch, _lib, L, A, b, oc, stream, 200)
        eq["n512_matrix_free_vs_fp64_csr"] = eq_matfree_n512(torch, airms, cf, sg, A, 100)
        eq["c3_batched_fp32"] = eq_batch(torch, _lib, L, airms, cf, sg, stream, 300, 256)

    cpu = None
    if not a.no_cpu:
        # the reference on the host: SART on up to 16 threads and on one, gradient descent, and
        # its serial matrix build, at N=512; SART at N=256 for the batched problem
        d2 = run_cpu_baseline(None, ("sart@16", "sart@1", "gd@16"), time_build=True, iters=a.cpu_iters)
        s16, s1, gd16 = leg(d2, 0), leg(d2, 1), leg(d2, 2)
        if s16 is not None:
            cpu = {"value": round(s16["iter_per_s"], 4), "unit": "iter/s", "cores": s16["cores"],
                   "kind": d2["kind"], "sample": s16["sample"] + f"; {d2['cpu_model']}",
                   "gbs": None if s16["gbs"] is None else round(s16["gbs"], 2),
                   "nproc": d2["nproc"], "affinity": d2["affinity"],
                   "value_1thread": None if s1 is None else round(s1["iter_per_s"], 4),
                   "reference_build_s": (None if d2.get("reference_build_s") is None
                                         else round(d2["reference_build_s"], 3))}
        if eq is not None:
            e = eq["c2_fp32_vs_fp64"]
            if s16 is not None:
                e["cpu_reference_s"] = round(e["iterations"] / s16["iter_per_s"], 3)
                e["cpu_cores"] = s16["cores"]
            if s1 is not None:
                e["cpu_reference_1thread_s"] = round(e["iterations"] / s1["iter_per_s"], 3)
            e = eq["n512_matrix_free_vs_fp64_csr"]
            if gd16 is not None:
                e["cpu_reference_s"] = round(e["iterations"] / gd16["iter_per_s"], 3)
                e["cpu_cores"] = gd16["cores"]
            d3 = run_cpu_baseline(cf.C3, ("sart@16",), iters=a.cpu_iters)
            s3 = leg(d3, 0)
            e = eq["c3_batched_fp32"]
            if s3 is not None:
                e["cpu_reference_s"] = round(e["slices"] * e["iterations"] / s3["iter_per_s"], 2)
                e["cpu_cores"] = s3["cores"]
                e["cpu_sample"] = s3["sample"]

    out = {
        "metric": METRIC, "value": round(it_s, 3), "unit": "iter/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(t_sart / a.steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded grains phantom, 0.5% seeded Gaussian noise)",
        "config": {"workload": "N=512 fan beam, 720 angles x 725 rays, SART FP64, box [0,1]",
                   "n": g.n, "angles": g.nth, "rays": g.p, "rows": m, "cols": n, "nnz": nnz,
                   "global_batch": 1, "seq_len": 0, "parallelism": "single GPU",
                   "sart_gbs_per_iteration": round(bytes_it / (t_sart / a.steps) / 1e9, 2),
                   "pcie_inclusive_iter_per_s": round(a.steps / t_host, 3),
                   "matrix_build_s": round(t_build, 4), "matrix_build_first_call_s": round(t_build_first, 4)},
        "roofline": {"kernel": "k_spmv_blocked<double, SrcPtr<double>, EpiStore<double>>",
                     "bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "avg_launch_us": round(t_spmv * 1e6, 3), "algorithmic_bytes": bytes_spmv},
        "cpu_baseline": cpu,
        "equal_error": eq,
    }
    print(json.dumps(out), flush=True)
    return 0


def _events(torch, stream):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def bench_precision(a, np, torch, _lib, airms, cf, sg, L, stream) -> int:
    """FP32 (values, iterate, accumulation; FP64 weights) against the plain FP64 method on the
    N=512 problem at several iteration counts, so time to equal error (<= 1e-5) can be read off."""
    g = cf.C2
    A = sg.fanlineartomo(sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta()))
    b = cf.add_noise(A.matvecmul(cf.grains(g.n, 200, 2)), 0.005, 2)
    oc = airms.Opts().set_lbound(0.0).set_ubound(1.0)._c()
    _lib.check(L.fair_sirt_prepare(A.handle, 1))
    rows = [eq_precision(torch, _lib, L, A, b, oc, stream, k) for k in (25, 50, 100, 200)]
    cpu = None if a.no_cpu else cpu_baseline(a.cpu_iters)
    last = rows[-1]
    out = {"metric": METRIC, "mode": "precision", "unit": "s", "dtype": "f32 vs f64",
           "config": {"workload": "N=512 fan beam, 720 x 725 rays, SART, box [0,1]"}, "rows": rows,
           "time_to_equal_error": dict(last, reference_cpu_s=None if cpu is None
                                       else round(last["iterations"] / cpu["value"], 3)),
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def bench_batch(a, np, torch, _lib, airms, cf, sg, L, stream) -> int:
    """256 slices of 256^2 sharing one fan-beam matrix (360 angles x 363 rays): one batched FP32
    SIRT (SART weights) run captured in a HIP graph against the plain FP64 method slice by slice,
    with the reference's SART timed on the host (per slice, scaled to the batch)."""
    r = eq_batch(torch, _lib, L, airms, cf, sg, stream, a.batch_iters, 256)
    out = {"metric": METRIC, "mode": "batch", "unit": "slice-iterations/s", "dtype": "f32 (fp64 weights)",
           "value": round(r["slices"] * r["iterations"] / r["batched_fp32_s"], 2), **r}
    cpu = None if a.no_cpu else cpu_baseline(a.cpu_iters, geom=cf.C3)
    if cpu is not None:
        out["cpu_baseline"] = cpu
        out["reference_cpu_s"] = round(r["slices"] * r["iterations"] / cpu["value"], 2)
        log(f"reference on {cpu['cores']} host threads: {cpu['value']} slice-iterations/s, "
            f"{out['reference_cpu_s']:.1f} s for the whole batch")
    print(json.dumps(out), flush=True)
    return 0


def bench_matfree(a, np, torch, _lib, airms, cf, sg, L, stream) -> int:
    """Matrix-free FP32 projector: 2048^2 Shepp-Logan + seeded Gaussian blobs, fan beam 2048 angles
    x 2900 rays (about 1.0e10 ray-pixel segments per product; the CSR would exceed int32 nnz):
    forward / adjoint timings, normest, gradient descent (box [0,1], fixed iteration count), the
    dot-product check; then at N=512 (720 x 725 rays) gradient descent against the plain FP64 CSR
    method from the same sigma, for time to equal error."""
    from fortran_airtools_amd.matrix_free import FanProjector
    dev = torch.device("cuda:0")
    iters = a.mf_iters
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    g = cf.C4
    P = FanProjector(sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta()))
    nseg = P.segments()
    xt = torch.from_numpy(cf.blobs(g.n).astype(np.float32)).to(dev)
    b = torch.empty(P.m, dtype=torch.float32, device=dev)
    P.matvecmul_dev(xt.data_ptr(), b.data_ptr())
    tmp_x = torch.empty(P.n, dtype=torch.float32, device=dev)
    tmp_y = torch.empty(P.m, dtype=torch.float32, device=dev)
    e0, e1 = _events(torch, stream)
    times = {}
    for name, f in (("forward", lambda: P.matvecmul_dev(xt.data_ptr(), tmp_y.data_ptr())),
                    ("adjoint", lambda: P.transvecmul_dev(b.data_ptr(), tmp_x.data_ptr()))):
        f()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(5):
            f()
        e1.record(stream)
        torch.cuda.synchronize()
        times[name] = e0.elapsed_time(e1) / 5 / 1e3
    log(f"2048^2: {nseg:.4g} segments/product; forward {times['forward'] * 1e3:.2f} ms, "
        f"adjoint {times['adjoint'] * 1e3:.2f} ms")
    # dot-product check <A x, y> = <x, A^T y> on two uniform pairs
    gen = torch.Generator(device=dev)
    gen.manual_seed(4)
    dots = []
    for _ in range(2):
        u = torch.rand(P.n, generator=gen, device=dev)
        v = torch.rand(P.m, generator=gen, device=dev)
        P.matvecmul_dev(u.data_ptr(), tmp_y.data_ptr())
        P.transvecmul_dev(v.data_ptr(), tmp_x.data_ptr())
        lhs = float(torch.dot(tmp_y.double(), v.double()))
        rhs = float(torch.dot(u.double(), tmp_x.double()))
        dots.append(abs(lhs - rhs) / abs(lhs))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sigma = P.normest(v0=np.random.default_rng(4).random(P.n))
    t_norm = time.perf_counter() - t0
    x = torch.zeros(P.n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = P.gradient_descent_dev(b.data_ptr(), x.data_ptr(), max_iter=iters, tolerance=0.0, sigma=sigma, options=o)
    torch.cuda.synchronize()
    t_gd = time.perf_counter() - t0
    err_truth = float(torch.linalg.norm(x.double() - xt.double()) / torch.linalg.norm(xt.double()))
    log(f"normest {t_norm:.3f} s (sigma {sigma:.6g}); GD {its} iterations {t_gd:.3f} s, "
        f"rel. error to phantom {err_truth:.3e}")
    del tmp_x, tmp_y, x, b, xt
    P.close()
    # N=512: FP64 CSR gradient descent against matrix-free FP32 from the same sigma
    g2 = cf.C2
    A = sg.fanlineartomo(sg.Data(n=g2.n, p=g2.p, nth=g2.nth, r=g2.r, dw=g2.dw, sd=g2.sd, theta=g2.theta()))
    n512 = eq_matfree_n512(torch, airms, cf, sg, A, iters)
    cpu = None if a.no_cpu else cpu_baseline(a.cpu_iters, "gd")
    out = {"metric": METRIC, "mode": "matfree", "unit": "GD iter/s", "dtype": "f32 (fp64 ray geometry and norms)",
           "value": round(its / t_gd, 3),
           "config": {"workload": "2048x2048 Shepp-Logan + Gaussian blobs, fan beam 2048 x 2900 rays, matrix-free "
                                  "gradient descent, box [0,1]", "iterations": its,
                      "segments_per_product": nseg},
           "forward_ms": round(times["forward"] * 1e3, 3), "adjoint_ms": round(times["adjoint"] * 1e3, 3),
           "roofline": mf_roofline(nseg, times),
           "segments_per_s": {"forward": round(nseg / times["forward"], 1),
                              "adjoint": round(nseg / times["adjoint"], 1),
                              "gd": round(2 * nseg * its / t_gd, 1)},
           "normest_s": round(t_norm, 4), "sigma": sigma, "gd_s": round(t_gd, 4),
           "relerr_to_phantom": err_truth, "dot_product_relerr": max(dots),
           "time_to_equal_error_n512": n512,
           "cpu_baseline_n512_gd": cpu}
    print(json.dumps(out), flush=True)
    return 0


FP64_VECTOR_PEAK_TFLOPS = 78.6   # also the FP32 vector rate without packed FP32 instructions
VALU_LANE_OPS_PEAK = 256 * 64 * 2.4e9   # 256 CUs, 64 lanes a cycle, 2.4 GHz


def mf_roofline(nseg: float, times: dict) -> dict:
    """The matrix-free products against the vector peaks: floating-point rate (the counted flops
    per segment times the measured segments per second) and VALU issue (lane operations)."""
    c = matfree_counts()
    out = {"bound": "valu", "unit": "TFLOP/s", "peak": FP64_VECTOR_PEAK_TFLOPS,
           "valu_lane_ops_peak_per_s": VALU_LANE_OPS_PEAK}
    for k in ("forward", "adjoint"):
        fl, ops = c.get(f"{k}_fp_flops_per_segment"), c.get(f"{k}_valu_lane_ops_per_segment")
        if fl is None or ops is None:
            out[k] = None
            continue
        rate = nseg / times[k]
        out[k] = {"achieved": round(fl * rate / 1e12, 3), "frac": round(fl * rate / 1e12 / FP64_VECTOR_PEAK_TFLOPS, 4),
                  "valu_issue_frac": round(ops * rate / VALU_LANE_OPS_PEAK, 4),
                  "flops_per_segment": fl, "valu_lane_ops_per_segment": ops}
    return out


def bench_kaczmarz(a, np, torch, _lib, airms, cf, sg, L, stream) -> int:
    """Kaczmarz at N=256 (fan beam 360 x 363 rays, three-phase phantom, 1 % noise, seed 5), box
    [0,1], 20 sweeps: randomized Kaczmarz with row-norm sampling and with uniform (scrambled)
    order (the reference's random_number draws supplied from a seeded generator), symmetric
    Kaczmarz (K=20: 10 forward + reverse pairs) and cyclic Kaczmarz (K=20), twin_art with
    symmetric Kaczmarz and the errorgauge rule, and cyclic Kaczmarz (4 sweeps) on the N=512 SART
    problem.
    Each GPU result is checked against the oracle run on the same draws; the reference's own art
    and twin_art (one host thread, as they are serial) are timed on the same problem."""
    g = cf.C3
    dat = sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta())
    A = sg.fanlineartomo(dat)
    IA, JA, vA = A.components()
    xt = cf.threephases(g.n, seed=5)
    b = cf.add_noise(A.matvecmul(xt), 0.01, 5)
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    nonempty = int(np.count_nonzero(np.diff(IA) > 0))
    from oracle.oracle import Oracle, Reference, build_oracle
    build_oracle()
    orc = Oracle()
    rows = []
    for method, rowprob in (("randkaczmarz", True), ("randkaczmarz", False), ("symkaczmarz", False),
                            ("kaczmarz", False)):
        o.rowprobability = rowprob
        u = None
        if method == "randkaczmarz":
            u = np.random.default_rng(5).random(airms.art_uniform_count(method, A, [20], options=o))
        # warm-up: one iteration; for the cyclic orders it also makes the level schedule the timed
        # run reuses (cached in the matrix), so its time is reported as first_call_s
        k_warm = 2 if method == "symkaczmarz" else 1
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        airms.art(method, A, b, [k_warm], options=o, uniforms=u)
        t_first = time.perf_counter() - t0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        X = airms.art(method, A, b, [20], options=o, uniforms=u)
        t_gpu = time.perf_counter() - t0
        Xo, _ = orc.art(method, IA, JA, vA, A.n, b, [20], u=u, lb=0.0, ub=1.0, rowprob=rowprob)
        err = float(np.linalg.norm(X[0] - Xo[0]) / np.linalg.norm(Xo[0]))
        steps = 20 * nonempty
        t_ref = None
        if Reference.available() and not a.no_cpu:
            ref = Reference()
            ref.set(0, IA, JA, vA, A.n)
            ref.rng_seed(5)
            t0 = time.perf_counter()
            ref.art(0, method, b, [20], lb=0.0, ub=1.0, rowprob=rowprob)
            t_ref = time.perf_counter() - t0
        log(f"{method}{' (row norms)' if rowprob else ''}: GPU {t_gpu:.3f} s ({steps / t_gpu:.3g} rows/s), "
            f"rel. error to oracle {err:.2e}"
            + (f"; reference on one host thread {t_ref:.3f} s" if t_ref else ""))
        rows.append({"method": method, "rowprobability": rowprob, "sweeps": 20, "row_steps": steps,
                     "gpu_s": round(t_gpu, 4), "rows_per_s": round(steps / t_gpu, 1),
                     "nnz_per_s": round(20 * A.nnz / t_gpu, 1), "relerr_vs_oracle": err,
                     "within_1e-12": err <= 1e-12,
                     "first_call_s": round(t_first, 4), "first_call_sweeps": k_warm,
                     "reference_host_s": None if t_ref is None else round(t_ref, 4)})
    # twin_art, symmetric Kaczmarz with the errorgauge rule: two chains per iteration, up to 20
    # sweeps each (the rule may stop earlier; final_it says where)
    o.rowprobability = False
    airms.twin_art("symkaczmarz", "errorgauge", A, b, [2], options=o)          # warm-up (schedules)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xt_, fit, _, _, _ = airms.twin_art("symkaczmarz", "errorgauge", A, b, [20], options=o)
    t_gpu = time.perf_counter() - t0
    xo_, fio, _, _, _, _ = orc.twin_art("symkaczmarz", "errorgauge", IA, JA, vA, A.n, b, [20], lb=0.0, ub=1.0)
    err = float(np.linalg.norm(xt_ - xo_) / np.linalg.norm(xo_))
    t_ref = None
    if Reference.available() and not a.no_cpu:
        ref = Reference()
        ref.set(0, IA, JA, vA, A.n)
        t0 = time.perf_counter()
        ref.twin_art(0, "symkaczmarz", "errorgauge", b, [20], lb=0.0, ub=1.0)
        t_ref = time.perf_counter() - t0
    log(f"twin_art symkaczmarz errorgauge: GPU {t_gpu:.3f} s (final_it {fit}, oracle {fio}), rel. error to "
        f"oracle {err:.2e}" + (f"; reference on one host thread {t_ref:.3f} s" if t_ref else ""))
    rows.append({"method": "twin_art symkaczmarz errorgauge", "kin": 20, "final_it": fit, "final_it_oracle": fio,
                 "gpu_s": round(t_gpu, 4), "relerr_vs_oracle": err, "within_1e-12": err <= 1e-12,
                 "reference_host_s": None if t_ref is None else round(t_ref, 4)})
    # cyclic Kaczmarz on the N=512 SART problem (720 x 725 rays, rows up to about 1000 entries):
    # 4 sweeps against the oracle, the reference's sweep time from its own art on one host thread
    g2 = cf.C2
    A2 = sg.fanlineartomo(sg.Data(n=g2.n, p=g2.p, nth=g2.nth, r=g2.r, dw=g2.dw, sd=g2.sd, theta=g2.theta()))
    b2 = A2.matvecmul(cf.grains(g2.n, 200, 2))
    o.rowprobability = False
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    airms.art("kaczmarz", A2, b2, [1], options=o)                              # schedule + graph
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    X2 = airms.art("kaczmarz", A2, b2, [4], options=o)
    t_gpu = time.perf_counter() - t0
    I2, J2, v2 = A2.components()
    Xo2, _ = orc.art("kaczmarz", I2, J2, v2, A2.n, b2, [4], lb=0.0, ub=1.0)
    err = float(np.linalg.norm(X2[0] - Xo2[0]) / np.linalg.norm(Xo2[0]))
    ref2 = None if a.no_cpu else cpu_baseline(1, "kaczmarz", geom=g2)
    t_ref = None if ref2 is None else 4.0 / ref2["value"]
    log(f"kaczmarz N=512: GPU {t_gpu:.3f} s for 4 sweeps (first call, 1 sweep: {t_first:.3f} s), rel. error to "
        f"oracle {err:.2e}" + (f"; reference on one host thread {t_ref:.3f} s" if t_ref else ""))
    rows.append({"method": "kaczmarz", "n": g2.n, "sweeps": 4, "gpu_s": round(t_gpu, 4),
                 "first_call_s": round(t_first, 4), "first_call_sweeps": 1, "relerr_vs_oracle": err,
                 "within_1e-12": err <= 1e-12, "reference_host_s": None if t_ref is None else round(t_ref, 4),
                 "reference_sample": None if ref2 is None else ref2["sample"]})
    # uniform randomized Kaczmarz on the same problem (rows up to about 1000 entries: the window
    # sweep with 16 entries per lane), 2 sweeps; the reference's own art on one host thread
    u2 = np.random.default_rng(6).random(airms.art_uniform_count("randkaczmarz", A2, [2], options=o))
    airms.art("randkaczmarz", A2, b2, [1], options=o, uniforms=u2)                 # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    X3 = airms.art("randkaczmarz", A2, b2, [2], options=o, uniforms=u2)
    t_gpu = time.perf_counter() - t0
    Xo3, _ = orc.art("randkaczmarz", I2, J2, v2, A2.n, b2, [2], u=u2, lb=0.0, ub=1.0)
    err = float(np.linalg.norm(X3[0] - Xo3[0]) / np.linalg.norm(Xo3[0]))
    ref3 = None if a.no_cpu else cpu_baseline(1, "randkaczmarz", geom=g2)
    t_ref = None if ref3 is None else 2.0 / ref3["value"]
    log(f"randkaczmarz N=512: GPU {t_gpu:.3f} s for 2 sweeps, rel. error to oracle {err:.2e}"
        + (f"; reference on one host thread {t_ref:.3f} s" if t_ref else ""))
    rows.append({"method": "randkaczmarz", "n": g2.n, "sweeps": 2, "gpu_s": round(t_gpu, 4),
                 "relerr_vs_oracle": err, "within_1e-12": err <= 1e-12,
                 "reference_host_s": None if t_ref is None else round(t_ref, 4),
                 "reference_sample": None if ref3 is None else ref3["sample"]})
    out = {"metric": METRIC, "mode": "kaczmarz", "unit": "rows/s", "dtype": "f64",
           "config": {"workload": "N=256 fan beam 360 x 363 rays, Kaczmarz 20 sweeps, box [0,1]",
                      "nnz": A.nnz, "nonempty_rows": nonempty},
           "rows": rows}
    print(json.dumps(out), flush=True)
    return 0


def bench_phantoms(a, np, torch) -> int:
    """The phantom gallery (phantomgallery.f90) and add_white_noise (system_generation.f90:709-745)
    on the device: every phantom at N=2048 and N=256 (host-side draws and shape logic included,
    the runtime's random_number replaced by a seeded stream), the noise for a 2900 x 2048 and a
    400 x 384 sinogram.  The reference's own functions (oracle/_ref, one host thread, seed 1) are
    timed at N=256 and on the 400 x 384 sinogram."""
    from fortran_airtools_amd import phantomgallery as pg, system_generation as sg
    from oracle.oracle import Reference
    names = ["shepplogan", "smooth", "threephases", "threephases_smooth", "binary", "fourphases", "mst",
             "grains", "bubbles"]

    def gpu(name, n):
        fn = getattr(pg, name)
        args = {} if name in ("shepplogan", "smooth") else {"draws": pg.Draws(seed=1)}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        img = fn(n, **args)
        t = time.perf_counter() - t0
        assert img.shape == (n, n) and np.isfinite(img).all()
        return t

    ref = Reference() if (Reference.available() and not a.no_cpu) else None
    rows = []
    for name in names:
        gpu(name, 64)                                    # warm-up (kernel loads)
        t2048, t256 = gpu(name, 2048), gpu(name, 256)
        t_ref = None
        if ref is not None:
            t0 = time.perf_counter()
            if name == "shepplogan":
                ref.shepplogan(256)
            else:
                ref.phantom(name, 256, 1)
            t_ref = time.perf_counter() - t0
        log(f"{name}: device {t2048 * 1e3:.1f} ms at N=2048, {t256 * 1e3:.1f} ms at N=256"
            + (f"; reference {t_ref * 1e3:.1f} ms at N=256" if t_ref is not None else ""))
        rows.append({"phantom": name, "gpu_n2048_s": round(t2048, 5), "gpu_n256_s": round(t256, 5),
                     "reference_n256_s": None if t_ref is None else round(t_ref, 5)})
    noise = []
    for m in (2900 * 2048, 400 * 384):
        rng = np.random.default_rng(2)
        b, u = rng.random(m), rng.random(m)
        sg.add_white_noise(b[:1000], 0.01, uniforms=u[:1000])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sg.add_white_noise(b, 0.01, uniforms=u)
        t = time.perf_counter() - t0
        t_ref = None
        if ref is not None and m == 400 * 384:
            t0 = time.perf_counter()
            ref.white_noise(b, 0.01, 1)                  # seed 1: the reference's Newton converges
            t_ref = time.perf_counter() - t0
        log(f"add_white_noise, {m} samples: device {t * 1e3:.1f} ms"
            + (f"; reference {t_ref * 1e3:.1f} ms" if t_ref is not None else ""))
        noise.append({"samples": m, "gpu_s": round(t, 5), "reference_s": None if t_ref is None else round(t_ref, 5)})
    out = {"metric": METRIC, "mode": "phantoms", "unit": "s", "higher_is_better": False, "dtype": "f64",
           "config": {"workload": "phantom gallery at N=2048 / N=256 and add_white_noise (seeded draws)"},
           "rows": rows, "noise": noise}
    print(json.dumps(out), flush=True)
    return 0


def bench_uneven(a, np, torch, _lib, airms, cf, sg, L, stream) -> int:
    """The uneven batch: 64 seeded SART problems (N from 128 to 1024, or the --uneven-sizes set;
    N/2..2N angles, gallery phantoms, 1 % noise, box [0,1]), each with its own device-built matrix,
    iterated 20 times as one group (two persistent launches per iteration for all problems) against
    the same problems run one after another; the group's iterates must equal the single runs bit
    for bit."""
    dev = torch.device("cuda:0")
    iters = 20
    sizes = [int(v) for v in a.uneven_sizes.split(",")] if a.uneven_sizes else None
    probs = cf.c5_problems(a.uneven_count, 5, sizes)
    gens = [cf.shepplogan, cf.bubbles, cf.smooth, cf.threephases]
    o = airms.Opts().set_lbound(0.0).set_ubound(1.0)
    mats, db = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k, g in enumerate(probs):
        mats.append(sg.fanlineartomo(sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta())))
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    for k, (g, A) in enumerate(zip(probs, mats)):
        x = gens[k % 4](g.n, seed=5 + k) if gens[k % 4] is not cf.shepplogan else cf.shepplogan(g.n)
        db.append(torch.from_numpy(cf.add_noise(A.matvecmul(x), 0.01, 5 + k)).to(dev))
    nnz = sum(A.nnz for A in mats)
    log(f"{len(mats)} problems, N {min(g.n for g in probs)}..{max(g.n for g in probs)}, total nnz {nnz:.4g}, "
        f"device builds {t_build:.2f} s")
    # staged-layout shapes per matrix (untimed set-up; the single runs below use the same ones)
    t0 = time.perf_counter()
    ch = (C.c_int32 * 6)()
    for A in mats:
        _lib.check(L.fair_autotune_spmv(A.handle, ch))
    t_tune = time.perf_counter() - t0
    log(f"staged-layout autotuning of {len(mats)} matrices: {t_tune:.2f} s")
    t0 = time.perf_counter()
    G = airms.SirtGroup(mats, "sart")
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    wg = G.autotune()
    xg = [torch.zeros(A.n, dtype=torch.float64, device=dev) for A in mats]
    G.sirt_dev([b.data_ptr() for b in db], 1, [x.data_ptr() for x in xg], o)     # warm-up
    for x in xg:
        x.zero_()
    torch.cuda.synchronize()
    e0, e1 = _events(torch, stream)
    e0.record(stream)
    G.sirt_dev([b.data_ptr() for b in db], iters, [x.data_ptr() for x in xg], o)
    e1.record(stream)
    torch.cuda.synchronize()
    t_group = e0.elapsed_time(e1) / 1e3
    grid = "one workgroup per item" if wg == 0 else f"{wg} workgroups per CU"
    log(f"group: {iters} iterations {t_group * 1e3:.1f} ms ({grid}), set-up {t_setup:.2f} s")
    xs = [torch.zeros(A.n, dtype=torch.float64, device=dev) for A in mats]
    for A, b, x in zip(mats, db, xs):       # weights, layouts and workspaces, untimed
        airms.sirt_dev("sart", A, b.data_ptr(), 1, x.data_ptr(), o)
        x.zero_()
    torch.cuda.synchronize()
    e0.record(stream)
    for A, b, x in zip(mats, db, xs):
        airms.sirt_dev("sart", A, b.data_ptr(), iters, x.data_ptr(), o)
    e1.record(stream)
    torch.cuda.synchronize()
    t_seq = e0.elapsed_time(e1) / 1e3
    same = all(bool(torch.equal(u, v)) for u, v in zip(xg, xs))
    log(f"one problem after another: {t_seq * 1e3:.1f} ms; group iterates bit-identical: {same}")
    bytes_it = sum(24 * A.nnz + 4 * (A.m + A.n + 2) + 32 * A.m + 32 * A.n for A in mats)
    out = {"metric": METRIC, "mode": "uneven", "unit": "problem-iterations/s", "dtype": "f64",
           "value": round(len(mats) * iters / t_group, 2),
           "config": {"workload": f"{len(mats)} fan-beam SART problems, N {min(g.n for g in probs)}.."
                                  f"{max(g.n for g in probs)}, N/2..2N angles, box [0,1], "
                                  f"{iters} iterations, one group", "total_nnz": nnz,
                      "sizes": [g.n for g in probs]},
           "group_s": round(t_group, 5), "sequential_s": round(t_seq, 5), "speedup": round(t_seq / t_group, 3),
           "group_gbs": round(bytes_it * iters / t_group / 1e9, 1), "workgroups_per_cu": wg,
           "bit_identical_to_single_runs": same, "device_build_s": round(t_build, 3),
           "layout_autotune_s": round(t_tune, 2),
           "group_setup_s": round(t_setup, 3)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
