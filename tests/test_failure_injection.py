"""Failure paths of the library's device allocations, on the CPU.

Every device allocation goes through one hook; FAIR_FAIL_ALLOC=N makes the N-th allocation of the
process fail as an out-of-memory hipMalloc would, before hipMalloc is called.  The entry point that
hit it must return FAIR_ERR_HIP with the allocation's message and leave nothing half-made: the
matrix handle is not returned.  Without the switch, a process with no GPU must fail just as loudly
(the product path has no fallback).  The Fortran drop-in turns such a status into the reference's
`stop "<message>"` (tests/test_gpu_fortran_dropin.py runs that on the GPU).  Each case runs in its
own child process (the allocation count is process-wide)."""
import json
import os
import subprocess
import sys

from conftest import ROOT

CHILD = r"""
import ctypes as C, json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from fortran_airtools_amd import _lib
L = _lib.load()
IA = np.array([0, 2, 3], np.int32)
JA = np.array([1, 3, 2], np.int32)
vA = np.array([1.0, 2.0, 3.0])
h = C.c_void_p()
st = L.fair_csr_create(2, 3, 3, IA.ctypes.data_as(C.c_void_p), JA.ctypes.data_as(C.c_void_p),
                       vA.ctypes.data_as(C.c_void_p), C.byref(h))
print(json.dumps({"status": st, "message": L.fair_last_error().decode(), "handle": h.value}))
"""


def child(**env):
    e = dict(os.environ, FAIR_NO_TORCH="1")
    e.pop("FAIR_FAIL_ALLOC", None)
    e.update(env)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True, timeout=120, env=e)
    assert r.returncode == 0, r.stdout + r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_injected_allocation_failure():
    from fortran_airtools_amd import _lib
    d = child(FAIR_FAIL_ALLOC="1")
    assert d["status"] == _lib.FAIR_ERR_HIP
    assert "out of memory" in d["message"] and "allocation 1 failed by FAIR_FAIL_ALLOC" in d["message"]
    assert d["handle"] is None


def test_unmatched_injection_does_not_fire_and_no_device_fails_loudly():
    """With the switch set past the allocations made, the call proceeds to hipMalloc: with no GPU
    visible (this test hides the devices) that fails with HIP's own error, not the injected one."""
    from fortran_airtools_amd import _lib
    d = child(FAIR_FAIL_ALLOC="1000", HIP_VISIBLE_DEVICES="-1")
    assert d["status"] == _lib.FAIR_ERR_HIP
    assert "FAIR_FAIL_ALLOC" not in d["message"] and "hipMalloc" in d["message"]
    assert d["handle"] is None
