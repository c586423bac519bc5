"""CPU tests of the product's host side: the per-cell kernel code compiled for the host
(tests/hostcheck) against the oracle bit for bit, the C ABI library's exports, and the Newton
control template with the reference's scalar mock."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle_ffi as O
from tests.helpers import ROOT, bits, dp, ensure_oracle, hostcheck, ip, perturbed_state


@pytest.fixture(scope="module", autouse=True)
def _libs():
    ensure_oracle()
    hostcheck()


def product_problem(po):
    H = hostcheck()
    P2 = O.Problem()
    K = np.empty(po.N)
    bt = np.empty(po.nbf, np.int32)
    bv = np.empty(po.nbf * po.m)
    u0 = np.empty(po.N * po.m)
    assert H.hc_build(C.byref(po.spec), C.byref(P2), dp(K), ip(bt), dp(bv), dp(u0)) == 0
    return P2, K, bt, bv, u0


CASES = [(1, 2, (12, 9), {}), (1, 3, (5, 6, 4), dict(gravity=0)), (2, 3, (8, 7, 6), {}),
         (2, 2, (16, 12), {}), (2, 3, (1, 5, 7), {}), (2, 3, (6, 6, 6), dict(gravity=0))]


@pytest.mark.parametrize("kind,dim,n,kw", CASES)
def test_builder_material_pattern_bit_exact(kind, dim, n, kw):
    po = O.OracleProblem(kind=kind, dim=dim, n=n, seed=5, **kw)
    P2, K, bt, bv, u0 = product_problem(po)
    assert np.array_equal(bits(K), bits(po.K))
    assert np.array_equal(bt, po.bc_type) and np.array_equal(bits(bv), bits(po.bc_value))
    assert np.array_equal(bits(u0), bits(po.u0))
    if kind == 2:
        mp = np.empty(24)
        hostcheck().hc_material_params(C.byref(P2), dp(mp))
        assert np.array_equal(bits(mp), bits(po.material_params()))
    rp, ci = po.pattern()
    H = hostcheck()
    nnz = H.hc_pattern(C.byref(P2), None, None)
    rp2, ci2 = np.empty(po.N + 1, np.int32), np.empty(nnz, np.int32)
    H.hc_pattern(C.byref(P2), ip(rp2), ip(ci2))
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2)


@pytest.mark.parametrize("kind,dim,n,kw", CASES)
def test_host_compiled_kernels_bit_exact(kind, dim, n, kw):
    po = O.OracleProblem(kind=kind, dim=dim, n=n, seed=6, **kw)
    P2, *_ = product_problem(po)
    u = perturbed_state(po, po.u0, seed=2)
    uprev = perturbed_state(po, po.u0, seed=3)
    r1, J1 = po.assemble(u, uprev, 10.0)
    r2, J2 = np.empty_like(r1), np.empty_like(J1)
    hostcheck().hc_assemble(C.byref(P2), dp(u), dp(uprev), 10.0, dp(r2), dp(J2))
    assert np.array_equal(bits(r1), bits(r2))
    assert np.array_equal(bits(J1), bits(J2))


def test_level_schedule_definition():
    po = O.OracleProblem(kind=2, dim=3, n=(5, 4, 3))
    rp, ci = po.pattern()
    lp = np.empty(po.N + 1, np.int32)
    lr = np.empty(po.N, np.int32)
    nl = hostcheck().hc_level_schedule(po.N, ip(rp), ip(ci), ip(lp), ip(lr))
    assert nl == 5 + 4 + 3 - 2               # wavefronts i + j + k = const of the 7-point stencil
    lvl = np.empty(po.N, int)
    for l in range(nl):
        lvl[lr[lp[l]:lp[l + 1]]] = l
    i, j, k = np.arange(po.N) % 5, (np.arange(po.N) // 5) % 4, np.arange(po.N) // 20
    assert np.array_equal(lvl, i + j + k)
    for l in range(nl):
        assert np.all(np.diff(lr[lp[l]:lp[l + 1]]) > 0)


def test_c_abi_exports_every_declared_symbol():
    lib_path = os.path.join(ROOT, "dumux_amd", "build", "libdumux_amd.so")
    if not os.path.exists(lib_path):
        from dumux_amd import build as B
        B.build(verbose=False)
    header = open(os.path.join(ROOT, "include", "dumux_amd.h")).read()
    declared = set(re.findall(r"\b(dmx_[a-z0-9_]+)\s*\(", header))
    declared -= {"dmx_context"}
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = declared - exported
    assert not missing, missing
    import dumux_amd as dx
    assert set(dx.EXPORTS) == declared
    L = C.CDLL(lib_path)     # loads without a GPU; no compute call is made
    for name in declared:
        assert hasattr(L, name)


def test_newton_template_with_reference_mock(tmp_path):
    exe = tmp_path / "newton_mock"
    src = os.path.join(ROOT, "tests", "newtonmock", "newton_mock.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), src], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ok" in p.stdout
