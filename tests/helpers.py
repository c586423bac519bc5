"""Shared test helpers: the oracle binding, the host-compiled product kernels (hostcheck)
and small-problem builders."""
import ctypes as C
import os
import subprocess

import numpy as np

from oracle import oracle_ffi as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOSTCHECK = os.path.join(ROOT, "tests", "hostcheck", "build", "libhostcheck.so")


def ensure_oracle():
    if not os.path.exists(O.LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    return O.lib()


_hc = None


def hostcheck():
    """The product's kernel headers compiled for the host (tests/hostcheck)."""
    global _hc
    if _hc is None:
        if not os.path.exists(HOSTCHECK):
            subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "hostcheck")], check=True, capture_output=True)
        L = C.CDLL(HOSTCHECK)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.hc_dpow.restype = C.c_double
        L.hc_dpow.argtypes = [C.c_double, C.c_double]
        L.hc_dlog.restype = C.c_double
        L.hc_dlog.argtypes = [C.c_double]
        L.hc_philox.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.hc_build.argtypes = [C.POINTER(O.BuildSpec), C.POINTER(O.Problem), dp, ip, dp, dp]
        L.hc_material_params.argtypes = [C.POINTER(O.Problem), dp]
        L.hc_pattern.restype = C.c_int64
        L.hc_pattern.argtypes = [C.POINTER(O.Problem), ip, ip]
        L.hc_level_schedule.restype = C.c_int
        L.hc_level_schedule.argtypes = [C.c_int64, ip, ip, ip, ip]
        L.hc_assemble.argtypes = [C.POINTER(O.Problem), dp, dp, C.c_double, dp, dp]
        fp = C.POINTER(C.c_float)
        L.hc_stencil.argtypes = [C.POINTER(O.Problem), dp, dp, C.c_double, fp, fp, fp, fp, fp]
        _hc = L
    return _hc


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def perturbed_state(p, u, seed=1, sn_max=0.3):
    """A state with non-trivial saturations (all regularization branches get exercised)."""
    rng = np.random.default_rng(seed)
    v = np.array(u, dtype=np.float64, copy=True)
    if p.m == 2:
        v[0::2] += rng.normal(0, 200.0, p.N)
        v[1::2] = np.clip(v[1::2] + rng.uniform(0, sn_max, p.N), 0.0, 0.999)
        # a few cells exactly at the regularization thresholds and beyond
        k = min(8, p.N)
        v[1:2 * k:2] = [0.0, 1e-12, 0.05, 0.5, 0.94, 0.95, 0.9505, 1.0][:k]
    else:
        v += rng.normal(0, 1e3, p.N)
    return v
