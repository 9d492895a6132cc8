"""CPU baseline timing of the reference path (test infrastructure; bench.py's cpu_baseline leg).

Times the reference's own `sirt`, `gradient_descent` (--method gd) or cyclic `art` (--method
kaczmarz, one host thread as the method is serial; its automatic arrays need a stack above 8 MB at
N=512, which the caller raises) (oracle/_ref, the reference's Fortran compiled with flang
-O3 -fopenmp; kind "reference") on the host cores, or the C restatement (kind "port",
single-threaded) when the compiled reference is absent.  The matrix comes from the C oracle's
ray tracer (bit-identical to the reference's fanlineartomo) so that the bounded sample does not
spend minutes in the reference's serial build.  The timed quantity is one iteration:
(t(k iterations) - t(0 iterations)) / k, so the serial set-up (row norms, double transpose;
gradient descent's normest, started from the same seed every run) cancels; k grows until the
iterations take at least 2 s.

usage: python -m oracle.cpu_baseline --n 512 --p 725 --nth 720 --step 0.5 --iters 3 --threads 16
Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--p", type=int, default=725)
    ap.add_argument("--nth", type=int, default=720)
    ap.add_argument("--step", type=float, default=0.5)
    ap.add_argument("--r", type=float, default=2.0)
    ap.add_argument("--dw", type=float, default=3.1)
    ap.add_argument("--sd", type=float, default=4.0)
    ap.add_argument("--method", default="sart")
    ap.add_argument("--iters", type=int, default=3)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    os.environ["OMP_NUM_THREADS"] = str(a.threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.dirname(here))
    import numpy as np
    from oracle.oracle import Oracle, Reference

    o = Oracle()
    th = (np.arange(a.nth) * a.step) / 180.0 * np.pi
    t0 = time.perf_counter()
    IA, JA, vA = o.fanlineartomo(a.n, a.p, a.nth, a.r, a.dw, a.sd, th)
    t_build = time.perf_counter() - t0
    n = a.n * a.n
    rng = np.random.default_rng(2)
    b = o.matvec(IA, JA, vA, rng.random(n))
    gd = a.method == "gd"
    kz = a.method == "kaczmarz"
    if kz and not Reference.available():
        print(json.dumps(None))
        return
    if Reference.available():
        ref = Reference()
        kind, cores = "reference", 1 if kz else a.threads

        def run(k):
            ref.set(0, IA, JA, vA, n)
            t = time.perf_counter()
            if kz:                       # k + 1 sweeps: the first carries the set-up
                ref.art(0, "kaczmarz", b, [k + 1], lb=0.0, ub=1.0)
            elif gd:
                ref.rng_seed(1)          # the same normest start (and iteration count) every run
                ref.gd(0, b, k, tol=0.0, lb=0.0, ub=1.0)
            else:
                ref.sirt(0, a.method, b, k, lb=0.0, ub=1.0)
            return time.perf_counter() - t
    else:
        kind, cores = "port", 1
        sigma = o.normest(IA, JA, vA, n, rng.random(n))[0] if gd else 0.0

        def run(k):
            t = time.perf_counter()
            if gd:
                o.gd(IA, JA, vA, n, b, k, sigma, tol=0.0, lb=0.0, ub=1.0)
            else:
                o.sirt(a.method, IA, JA, vA, n, b, k, lb=0.0, ub=1.0)
            return time.perf_counter() - t
    # set-up only (maxits = 0), then grow k until the iterations dominate the timing noise
    t_setup = sorted(run(0) for _ in range(3))[1]
    k = a.iters
    while True:
        t_k = run(k)
        if t_k - t_setup >= 2.0 or k >= 4096 or t_k > 30.0:
            break
        k *= 4
    per_iter = max(t_k - t_setup, 1e-9) / k
    bytes_iter = 24 * len(JA) + 4 * (len(IA) - 1 + n + 2) + 32 * (len(IA) - 1) + 32 * n
    print(json.dumps({"kind": kind, "cores": cores, "iter_per_s": 1.0 / per_iter,
                      "sec_per_iter": per_iter, "gbs": None if kz else bytes_iter / per_iter / 1e9,
                      "nnz": int(len(JA)), "oracle_build_s": t_build,
                      "setup_s": t_setup,
                      "sample": f"{a.method} on N={a.n} fan {a.nth}x{a.p}: "
                                f"(t({k} it) - t(0 it)) / {k}, {kind} on {cores} thread(s)"
                                + (" (an iteration is one sweep; t(k) runs k + 1 sweeps)" if kz else "")}))


if __name__ == "__main__":
    main()
