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


def cpu_baseline(iters: int) -> dict | None:
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--iters", str(iters), "--threads", str(cores)]
    log("cpu baseline:", " ".join(cmd))
    import resource

    def big_stack():
        # the reference keeps m- and n-sized automatic arrays on the stack (AIRMs.f90:1009-1010)
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    env = dict(os.environ, OMP_STACKSIZE="256M")
    try:
        r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=600,
                           preexec_fn=big_stack, env=env)
    except subprocess.TimeoutExpired:
        log("cpu baseline timed out")
        return None
    if r.returncode != 0:
        log(f"cpu baseline failed (status {r.returncode}):", r.stderr[-2000:])
        return None
    d = json.loads(r.stdout.strip().splitlines()[-1])
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": round(d["iter_per_s"], 4), "unit": "iter/s", "cores": d["cores"],
            "kind": d["kind"], "sample": d["sample"] + f"; {cpu_model}",
            "gbs": round(d["gbs"], 2)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 1 GiB fill and a 1 GiB sum (reference kernels for HBM counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              f"(asked for --gpus {a.gpus}).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from fortran_airtools_amd import _lib, airms, configs as cf, system_generation as sg

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    L = _lib.lib()
    stream = torch.cuda.current_stream()
    _lib.check(L.fair_set_stream(C.c_void_p(stream.cuda_stream)))

    g = cf.C2
    dat = sg.Data(n=g.n, p=g.p, nth=g.nth, r=g.r, dw=g.dw, sd=g.sd, theta=g.theta())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A = sg.fanlineartomo(dat)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    m, n, nnz = A.m, A.n, A.nnz
    log(f"matrix {m} x {n}, nnz {nnz}, device build {t_build:.3f} s")

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

    # the SpMV kernel alone (MMRS over the non-empty rows), HIP events on the launch stream
    xs = torch.rand(n, dtype=torch.float64, device=dev)
    ys = torch.zeros(m, dtype=torch.float64, device=dev)

    def spmv():
        _lib.check(L.fair_spmv_rows_dev(A.handle, C.c_void_p(xs.data_ptr()), C.c_void_p(ys.data_ptr())))

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

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(a.cpu_iters)

    out = {
        "metric": METRIC, "value": round(it_s, 3), "unit": "iter/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(t_sart / a.steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded grains phantom, 0.5% seeded Gaussian noise)",
        "config": {"workload": "N=512 fan beam, 720 angles x 725 rays, SART FP64, box [0,1]",
                   "n": g.n, "angles": g.nth, "rays": g.p, "rows": m, "cols": n, "nnz": nnz,
                   "global_batch": 1, "seq_len": 0, "parallelism": "single GPU",
                   "sart_gbs_per_iteration": round(bytes_it / (t_sart / a.steps) / 1e9, 2),
                   "matrix_build_s": round(t_build, 4)},
        "roofline": {"kernel": "k_spmv_blocked<double, SrcPtr<double>, EpiStore<double>>",
                     "bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "avg_launch_us": round(t_spmv * 1e6, 3), "algorithmic_bytes": bytes_spmv},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
