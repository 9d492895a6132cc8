"""CPU baseline timing of the reference path (test infrastructure; bench.py's cpu_baseline leg).

Times the reference's own `sirt` ('sart' / 'cimmino'), `gradient_descent` ('gd'), or `art`, cyclic
('kaczmarz') or uniform randomized ('randkaczmarz', the runtime's own draws), serial by construction (oracle/_ref, the reference's Fortran compiled with flang
-O3 -fopenmp; kind "reference") on the host cores, each leg with its own OpenMP thread count, or
the C restatement (kind "port", single-threaded) when the compiled reference is absent.  Its
automatic arrays need a stack above 8 MB at N=512, which the caller raises.

The matrix comes from the C oracle's ray tracer (bit-identical to the reference's
fanlineartomo) so that the bounded samples do not each spend the reference's serial build;
with --time-build the reference's own fanlineartomo is timed once as well (one host thread, as
it is serial).  The timed quantity of a leg is one iteration: (t(k iterations) - t(0)) / k, so
the serial set-up (row norms, double transpose; gradient descent's normest, started from the
same seed every run) cancels; k grows until the iterations take at least 2 s.  Progress goes to
stderr.  OpenMP threads are left unbound (OMP_PROC_BIND=false).

usage: python -m oracle.cpu_baseline --n 512 --p 725 --nth 720 --step 0.5 --legs sart@16,sart@1 [--time-build]
Prints one JSON object: host facts, and per leg "method@threads" its iterations per second.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def say(msg: str) -> None:
    """Progress on stderr (the JSON result goes to stdout)."""
    print("[cpu]", msg, file=sys.stderr, flush=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--p", type=int, default=725)
    ap.add_argument("--nth", type=int, default=720)
    ap.add_argument("--step", type=float, default=0.5)
    ap.add_argument("--r", type=float, default=2.0)
    ap.add_argument("--dw", type=float, default=3.1)
    ap.add_argument("--sd", type=float, default=4.0)
    ap.add_argument("--legs", default="sart@1", help="comma-separated method@threads")
    ap.add_argument("--iters", type=int, default=3, help="first k tried")
    ap.add_argument("--time-build", action="store_true")
    a = ap.parse_args()
    legs = []
    for leg in a.legs.split(","):
        meth, _, thr = leg.partition("@")
        legs.append((leg, meth, max(1, int(thr or 1))))
    os.environ["OMP_NUM_THREADS"] = str(max(t for _, _, t in legs))
    # no thread binding: on a one-GPU box (16 CPUs of a VM) OMP_PROC_BIND=close with
    # OMP_PLACES=cores put every thread of the team on one CPU (the reference's SART at N=256 ran
    # at 50 it/s on 1, 4 and 16 threads alike); unbound, 16 threads run 2.6x one
    os.environ["OMP_PROC_BIND"] = "false"
    os.environ.pop("OMP_PLACES", None)
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.dirname(here))
    import numpy as np
    from oracle.oracle import Oracle, Reference

    o = Oracle()
    th = (np.arange(a.nth) * a.step) / 180.0 * np.pi
    t0 = time.perf_counter()
    IA, JA, vA = o.fanlineartomo(a.n, a.p, a.nth, a.r, a.dw, a.sd, th)
    t_build = time.perf_counter() - t0
    say(f"oracle matrix N={a.n}: {t_build:.2f} s")
    n = a.n * a.n
    rng = np.random.default_rng(2)
    b = o.matvec(IA, JA, vA, rng.random(n))
    have_ref = Reference.available()
    ref = Reference() if have_ref else None
    out = {"kind": "reference" if have_ref else "port", "nproc": os.cpu_count(), "affinity": affinity,
           "cpu_model": cpu_model(), "n": a.n, "p": a.p, "nth": a.nth, "nnz": int(len(JA)),
           "oracle_build_s": t_build, "legs": {}}
    if a.time_build and have_ref:
        ref.set_threads(1)
        t0 = time.perf_counter()
        ref.build(0, 0, a.n, a.p, a.nth, a.r, a.dw, a.sd, th)
        out["reference_build_s"] = time.perf_counter() - t0
        say(f"reference fanlineartomo: {out['reference_build_s']:.2f} s")
    sigma = None
    for name, meth, thr in legs:
        gd, kz = meth == "gd", meth in ("kaczmarz", "randkaczmarz")
        if kz and not have_ref:
            out["legs"][name] = None
            continue
        if have_ref:
            cores = 1 if kz else thr
            ref.set_threads(cores)

            def run(k):
                ref.set(0, IA, JA, vA, n)
                t = time.perf_counter()
                if kz:                       # k + 1 sweeps: the first carries the set-up
                    ref.rng_seed(1)
                    ref.art(0, meth, b, [k + 1], lb=0.0, ub=1.0)
                elif gd:
                    ref.rng_seed(1)          # the same normest start (and iteration count) every run
                    ref.gd(0, b, k, tol=0.0, lb=0.0, ub=1.0)
                else:
                    ref.sirt(0, meth, b, k, lb=0.0, ub=1.0)
                return time.perf_counter() - t
        else:
            cores = 1
            if gd and sigma is None:
                sigma = o.normest(IA, JA, vA, n, rng.random(n))[0]

            def run(k):
                t = time.perf_counter()
                if gd:
                    o.gd(IA, JA, vA, n, b, k, sigma, tol=0.0, lb=0.0, ub=1.0)
                else:
                    o.sirt(meth, IA, JA, vA, n, b, k, lb=0.0, ub=1.0)
                return time.perf_counter() - t
        # set-up only (maxits = 0, the faster of two runs), then grow k until the iterations
        # dominate the timing noise
        t_setup = min(run(0), run(0))
        say(f"{name}: set-up {t_setup:.2f} s")
        k = a.iters
        while True:
            t_k = run(k)
            say(f"{name}: {k} iterations {t_k:.2f} s")
            if t_k - t_setup >= 2.0 or k >= 4096 or t_k > 30.0:
                break
            k *= 4
        per_iter = max(t_k - t_setup, 1e-9) / k
        bytes_iter = 24 * len(JA) + 4 * (len(IA) - 1 + n + 2) + 32 * (len(IA) - 1) + 32 * n
        out["legs"][name] = {
            "method": meth, "cores": cores, "iter_per_s": 1.0 / per_iter, "sec_per_iter": per_iter,
            "gbs": None if kz else bytes_iter / per_iter / 1e9, "setup_s": t_setup,
            "sample": f"{meth} on N={a.n} fan {a.nth}x{a.p}: (t({k} it) - t(0 it)) / {k}, "
                      f"{out['kind']} on {cores} thread(s)"
                      + (" (an iteration is one sweep; t(k) runs k + 1 sweeps)" if kz else "")}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
