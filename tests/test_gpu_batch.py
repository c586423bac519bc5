"""Batched executor (many problems, one set of launches per Newton phase) on the GPU.

Bars:
  * a problem's result in a batch is bitwise the same as the same problem alone in a batch,
    whatever the other problems are and whenever they join or leave (continuous batching);
  * against the CPU oracle: pw |a-b|/|b| <= 1e-8, Sn |a-b|/max(1,|b|) <= 1e-8.
"""
import numpy as np
import pytest

import dumux_amd as dx
from oracle import oracle_ffi as O
from tests.helpers import bits, ensure_oracle

pytestmark = pytest.mark.gpu

LIN = dict(reduction=1e-10, inner_reduction=1e-4, max_iter=1000, max_refinements=30)

# name: (kind, dim, n, builder kwargs, nsteps, dt, linear_pde); the stationary 1p problem is
# solved to 1e-13 (the oracle to 1e-14) so that two correct solutions agree to 1e-8 despite the
# system's conditioning
SPECS = {
    "2p3d": (2, 3, (12, 12, 12), dict(seed=21), 4, 10.0, False),
    "2p2d": (2, 2, (24, 20), dict(seed=4), 3, 10.0, False),
    "2p3d_b": (2, 3, (10, 9, 8), dict(seed=33, gravity=0), 2, 20.0, False),
    "1p2d": (1, 2, (40, 30), dict(seed=5, sigma_log10K=0.3), 1, 1.0, True),
    "1p3d": (1, 3, (9, 7, 5), dict(seed=2), 3, 50.0, False),
}


@pytest.fixture(scope="module", autouse=True)
def _libs():
    ensure_oracle()
    dx.lib()


def make(name):
    kind, dim, n, kw, nsteps, dt, lpde = SPECS[name]
    prob = dx.Problem(kind=kind, dim=dim, n=n, **kw)
    asm = dx.FVAssembler(prob)
    asm.setSolution(prob.u0)
    return prob, asm


def run_batch(names, late=(), leave=None, leave_after=2):
    """Run the named problems in one batch; `late` join after two steps, `leave` is removed
    after `leave_after` steps. Returns {name: (solution, finished, info)}."""
    b = dx.Batch(linear=LIN)
    members = {}
    for nm in names:
        prob, asm = make(nm)
        _, _, _, _, nsteps, dt, lpde = SPECS[nm]
        b.add(asm, nsteps=nsteps, dt=dt, linear_pde=lpde, reduction=1e-13 if lpde else 0.0)
        members[nm] = (prob, asm)
    steps = 0
    out = {}
    while True:
        if steps == 2:
            for nm in late:
                prob, asm = make(nm)
                _, _, _, _, nsteps, dt, lpde = SPECS[nm]
                b.add(asm, nsteps=nsteps, dt=dt, linear_pde=lpde, reduction=1e-13 if lpde else 0.0)
                members[nm] = (prob, asm)
        if leave is not None and steps == leave_after and leave in members:
            prob, asm = members.pop(leave)
            f, info = b.status(asm)
            out[leave] = (asm.solution(), f, info)
            b.remove(asm)
            asm.close()
        act = b.step()
        steps += 1
        if act == 0 and (steps > 2 or not late):
            break
        assert steps < 2000
    for nm, (prob, asm) in members.items():
        f, info = b.status(asm)
        out[nm] = (asm.solution(), f, info)
        asm.close()
    b.close()
    return out


def test_batch_matches_oracle():
    res = run_batch(list(SPECS))
    for nm, (u, f, info) in res.items():
        kind, dim, n, kw, nsteps, dt, lpde = SPECS[nm]
        assert f == 1, (nm, f)
        assert info.converged == 1
        po = O.OracleProblem(kind=kind, dim=dim, n=n, **kw)
        if lpde:
            u_o, info_o = po.newton(po.u0, None, 1.0, linear_pde=1, lin_reduction=1e-14, lin_max_iter=5000)
        else:
            u_o, info_o = po.time_loop(po.u0, nsteps, dt, lin_reduction=1e-10, lin_max_iter=1000)
            assert info.time_steps == nsteps
        assert info_o.converged
        m = po.m
        pw, pw_o = u[0::m], u_o[0::m]
        err = np.max(np.abs(pw - pw_o) / np.abs(pw_o))
        assert err <= 1e-8, (nm, err)
        if m == 2:
            sn_err = np.max(np.abs(u[1::2] - u_o[1::2]) / np.maximum(1.0, np.abs(u_o[1::2])))
            assert sn_err <= 1e-8, (nm, sn_err)


def test_batch_alone_equals_mixed_bitwise():
    alone = {nm: run_batch([nm])[nm] for nm in ("2p3d", "1p2d", "2p2d")}
    mixed = run_batch(list(SPECS))
    for nm, (u, f, info) in alone.items():
        u_m, f_m, info_m = mixed[nm]
        assert f == f_m == 1
        assert np.array_equal(bits(u), bits(u_m)), nm
        assert info.newton_steps == info_m.newton_steps
        assert info.linear_iterations == info_m.linear_iterations


def test_batch_continuous_join_and_leave():
    # 2p3d alone, against: others present from the start, one leaving early, two joining late
    ref = run_batch(["2p3d"])["2p3d"]
    res = run_batch(["2p3d", "2p2d", "1p3d"], late=("2p3d_b", "1p2d"), leave="2p2d", leave_after=3)
    u, f, info = res["2p3d"]
    assert f == 1
    assert np.array_equal(bits(u), bits(ref[0]))
    # the late joiners finished with their own results, equal to running alone
    for nm in ("2p3d_b", "1p2d"):
        assert res[nm][1] == 1
        assert np.array_equal(bits(res[nm][0]), bits(run_batch([nm])[nm][0])), nm
    # the one that left was still running (its solution is an intermediate iterate)
    assert res["2p2d"][1] == 0
