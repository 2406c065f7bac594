"""ctypes loader for the C restatement (oracle/isac_oracle.c).

TEST INFRASTRUCTURE ONLY (tests, smoke check and the CPU-baseline leg of
bench.py).  ``build()`` compiles it with the Makefile next to this file.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build() -> None:
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def _lib(precision: str = "f64"):
    name = {"f64": "liboracle64.so", "f32": "liboracle32.so"}[precision]
    if name not in _LIBS:
        path = os.path.join(_HERE, "_build", name)
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.oracle_mref.argtypes = [fp, fp, fp] + [ctypes.c_int] * 5 + \
            [ctypes.c_float] * 3 + [ctypes.c_int, fp, fp, ip, ip, dp, dp]
        lib.oracle_mref.restype = ctypes.c_int
        _LIBS[name] = lib
    return _LIBS[name]


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def mref(sbj, ref, acc_xy, ring_num, shift_rng_x, shift_rng_y, shift_step,
         ring_len=256, nthreads=0, precision="f64", tie_rel=1e-5):
    """Run the C oracle; returns a dict of [N, P] arrays like isac_oracle.multi_ref_align."""
    sbj = np.ascontiguousarray(sbj, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float32)
    P, D = sbj.shape[0], sbj.shape[-1]
    N = ref.shape[0]
    acc = np.ascontiguousarray(np.asarray(acc_xy, dtype=np.float32).reshape(P, 2))
    d = np.zeros((N, P), np.float32)
    peak = np.zeros((N, 4 * P), np.float32)
    sidx = np.zeros((N, P), np.int32)
    tidx = np.zeros((N, P), np.int32)
    delta = np.zeros((N, P), np.float64)
    runner = np.zeros((N, P), np.float64)
    rc = _lib(precision).oracle_mref(
        _ptr(sbj, ctypes.c_float), _ptr(ref, ctypes.c_float), _ptr(acc, ctypes.c_float),
        P, N, D, ring_num, ring_len, shift_rng_x, shift_rng_y, shift_step, nthreads,
        _ptr(d, ctypes.c_float), _ptr(peak, ctypes.c_float), _ptr(sidx, ctypes.c_int),
        _ptr(tidx, ctypes.c_int), _ptr(delta, ctypes.c_double), _ptr(runner, ctypes.c_double))
    if rc != 0:
        raise ValueError(f"oracle_mref rejected its arguments (code {rc})")
    best = d.astype(np.float64)
    return {"d": best, "angle": peak[:, 0::4].astype(np.float64),
            "sx": peak[:, 1::4].astype(np.float64), "sy": peak[:, 2::4].astype(np.float64),
            "mirror": peak[:, 3::4].astype(np.float64), "shift_idx": sidx, "t": tidx,
            "delta": delta, "runner": runner,
            "tie": (best - runner) <= tie_rel * np.abs(best), "peak_list": peak}
