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
        L.hc_mg_vcycle.restype = C.c_int
        L.hc_mg_vcycle.argtypes = [C.POINTER(O.Problem), dp, dp, C.c_double, fp, fp]
        L.hc_mg_vcycle_unscaled.restype = C.c_int
        L.hc_mg_vcycle_unscaled.argtypes = [C.POINTER(O.Problem), dp, dp, C.c_double, fp, fp]
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


# ---- the reference's 2p lens test (test/porousmediumflow/2p/incompressible) -------------------

LENS_FIELDS = ["S_aq", "p_aq", "S_napl", "p_napl", "pc", "mob_aq", "mob_napl"]


def lens_problem(p):
    """Turns a 2p 2D problem of n = (48, 32) cells (dumux_amd.Problem or oracle_ffi.OracleProblem)
    into test_2p_incompressible_tpfa's setup: domain (0,6)x(0,4) (params.input:5-8), porosity 0.4,
    lens (1,4)x(2,3) with K 9.05e-12 and van Genuchten Swr 0.18, alpha 4.5e-4, n 7.3, outside
    K 4.6e-10, Swr 0.05, alpha 3.7e-3, n 4.7 (spatialparams.hh:52-118, params.input:10-22);
    Dirichlet hydrostatic-like pw (Sn = 0) left and right, DNAPL influx 0.04 kg/(m s) on the top
    for 0.5 < (6-x)/6 < 2/3, no flow elsewhere, hydrostatic initial pw (problem.hh:66-150).
    Returns the material ids (1 = lens)."""
    d = p.desc if hasattr(p, "desc") else p.P
    nx, ny = 48, 32
    d.upper[0], d.upper[1] = 6.0, 4.0
    d.porosity = 0.4
    d.vg_alpha, d.vg_n, d.vg_l, d.swr, d.snr = 0.0037, 4.7, 0.5, 0.05, 0.0
    h = (6.0 / nx, 4.0 / ny)
    xc = np.array([i * h[0] + 0.5 * ((i + 1) * h[0] - i * h[0]) for i in range(nx)])
    yc = np.array([j * h[1] + 0.5 * ((j + 1) * h[1] - j * h[1]) for j in range(ny)])
    X, Y = (a.ravel() for a in np.meshgrid(xc, yc))
    eps = 1.5e-7
    lens = (X >= 1.0 + eps) & (X <= 4.0 - eps) & (Y >= 2.0 + eps) & (Y <= 3.0 - eps)
    bct = np.zeros(p.nbf, np.int32)
    bcv = np.zeros(p.nbf * 2)
    for j in range(ny):
        for side, x in ((0, 0.0), (1, 6.0)):
            b = side * ny + j
            alpha = 1 + 1.5 / 4.0
            factor = (6.0 * alpha + (1.0 - alpha) * x) / 6.0
            bct[b] = 1
            bcv[2 * b] = 1e5 - factor * 1000.0 * (-9.81) * (4.0 - yc[j])
    for i in range(nx):
        lam = (6.0 - xc[i]) / 6.0
        if 0.5 < lam < 2.0 / 3.0:
            bcv[2 * (2 * ny + nx + i) + 1] = -0.04
    p.set_arrays(K=np.where(lens, 9.05e-12, 4.6e-10), bc_type=bct, bc_value=bcv)
    mat = lens.astype(np.int32)
    p.set_materials(mat, vg_alpha=0.00045, vg_n=7.3, swr=0.18)
    p.u0[0::2] = 1e5 - 1000.0 * (-9.81) * (4.0 - Y)
    p.u0[1::2] = 0.0
    return mat


def lens_reference():
    """{field: values} of test/references/test_2p_incompressible_cc-reference.vtu (7th output)."""
    ref = np.loadtxt(os.path.join(ROOT, "tests", "golden", "test_2p_incompressible_cc_reference.txt"))
    return {nm: ref[:, k] for k, nm in enumerate(LENS_FIELDS)}


def fuzzy_close(a, ref, rel=1e-2, abs_scale=1.5e-7):
    """The reference's fuzzy VTU comparison (bin/testing/dumux_runtest.py:24-42, defaults :196-206):
    |a - b| <= max(abs_scale * max|ref|, rel * max(|a|, |b|)) for every value."""
    tol = np.maximum(abs_scale * np.abs(ref).max(), rel * np.maximum(np.abs(a), np.abs(ref)))
    return bool(np.all(np.abs(a - ref) <= tol))
