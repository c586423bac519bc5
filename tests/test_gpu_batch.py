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
LIN_MG = dict(LIN, precond=dx.PRECOND_AMG)   # aggregation multigrid preconditioner

# name: (kind, dim, n, builder kwargs, nsteps, dt, linear_pde); the stationary 1p problem is
# solved to 1e-13 (the oracle to 1e-14) so that two correct solutions agree to 1e-8 despite the
# system's conditioning
SPECS = {
    "2p3d": (2, 3, (12, 12, 12), dict(seed=21), 4, 10.0, False),
    "2p2d": (2, 2, (24, 20), dict(seed=4), 3, 10.0, False),
    "2p3d_b": (2, 3, (10, 9, 8), dict(seed=33, gravity=0), 2, 20.0, False),
    "1p2d": (1, 2, (40, 30), dict(seed=5, sigma_log10K=0.3), 1, 1.0, True),
    "1p3d": (1, 3, (9, 7, 5), dict(seed=2), 3, 50.0, False),
    # hand-coded derivatives (CCLocalAssembler<analytic>) next to the forward-difference members
    "2p3d_an": (2, 3, (11, 12, 10), dict(seed=22, diff_method=1), 3, 10.0, False),
    "1p2d_an": (1, 2, (36, 30), dict(seed=6, sigma_log10K=0.3, diff_method=1), 1, 1.0, True),
    # grids with several box-tiled multigrid levels before the one-workgroup tail (odd extents on
    # the way down)
    "2p3d_mg": (2, 3, (40, 36, 34), dict(seed=23, mu_log10K=-11.0, sigma_log10K=0.5), 2, 10.0, False),
    "2p2d_mg": (2, 2, (100, 90), dict(seed=8), 2, 10.0, False),
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


def run_batch(names, late=(), leave=None, leave_after=2, linear=LIN):
    """Run the named problems in one batch; `late` join after two ticks, `leave` is removed
    after `leave_after` ticks. Returns {name: (solution, finished, info)}."""
    b = dx.Batch(linear=linear)
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


@pytest.mark.parametrize("linear", [LIN, LIN_MG], ids=["block_jacobi", "multigrid"])
def test_batch_matches_oracle(linear):
    res = run_batch(list(SPECS), linear=linear)
    for nm, (u, f, info) in res.items():
        kind, dim, n, kw, nsteps, dt, lpde = SPECS[nm]
        assert f == 1, (nm, f)
        assert info.converged == 1
        assert info.fp64_fallback == 0, nm  # well-conditioned: FP32 refinement suffices
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


@pytest.mark.parametrize("linear", [LIN, LIN_MG], ids=["block_jacobi", "multigrid"])
def test_batch_alone_equals_mixed_bitwise(linear):
    alone = {nm: run_batch([nm], linear=linear)[nm] for nm in ("2p3d", "1p2d", "2p2d", "2p3d_an")}
    mixed = run_batch(list(SPECS), linear=linear)
    for nm, (u, f, info) in alone.items():
        u_m, f_m, info_m = mixed[nm]
        assert f == f_m == 1
        assert np.array_equal(bits(u), bits(u_m)), nm
        assert info.newton_steps == info_m.newton_steps
        assert info.linear_iterations == info_m.linear_iterations


def test_footprint_estimate_matches_allocation():
    for nm in ("2p3d", "1p2d", "2p3d_an"):
        prob, asm = make(nm)
        assert asm.device_bytes() == prob.estimate_bytes(batched=False)
        # in a batch: plus the batch arrays (and with multigrid its coarse levels)
        for lin, mg in ((LIN, False), (LIN_MG, True)):
            p2, a2 = make(nm)
            b = dx.Batch(linear=lin)
            b.add(a2, nsteps=1, dt=1.0)
            b.step()
            assert a2.device_bytes() == p2.estimate_bytes(batched=True, multigrid=mg), (nm, mg)
            b.close()
            a2.close()
        asm.close()


def test_multigrid_cuts_krylov_iterations():
    # aggregation multigrid against block Jacobi on the same problems: the same converged
    # results (<= 1e-8 apart) with a fraction of the FP32 BiCGSTAB iterations, and no FP64
    # fallback (a diverging coarse smoother would force one)
    names = ["2p3d", "2p3d_b", "1p2d", "2p3d_mg", "2p2d_mg"]
    bj = run_batch(names, linear=LIN)
    mg = run_batch(names, linear=LIN_MG)
    for nm in names:
        (u1, f1, i1), (u2, f2, i2) = bj[nm], mg[nm]
        assert f1 == f2 == 1
        factor = 3 if nm.endswith("_mg") else 2
        assert i2.linear_iterations * factor < i1.linear_iterations, (nm, i1.linear_iterations, i2.linear_iterations)
        assert i2.fp64_fallback == 0, nm
        m = SPECS[nm][0]
        assert np.max(np.abs(u1[0::m] - u2[0::m]) / np.abs(u1[0::m])) <= 1e-8


def test_batch_rerun_of_the_same_contexts_is_identical():
    # a second solve of the same contexts restarts from the solution set before adding them
    prob, asm = make("2p3d")
    out = []
    for rep in range(2):
        asm.setSolution(prob.u0)
        b = dx.Batch(linear=LIN)
        b.add(asm, nsteps=2, dt=10.0)
        b.run()
        f, info = b.status(asm)
        assert f == 1
        out.append((asm.solution(), info.newton_steps, info.linear_iterations))
        b.close()
    asm.close()
    assert np.array_equal(bits(out[0][0]), bits(out[1][0]))
    assert out[0][1:] == out[1][1:]


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


def test_served_stream_results_equal_alone():
    # Poisson arrivals, HBM budget that holds about two requests at once: requests wait,
    # join while others iterate, and leave as they finish; each result equals the request
    # solved alone (bitwise)
    from dumux_amd import serve
    rng = np.random.default_rng(7)
    reqs = serve.poisson_stream(rate=40.0, count=6, rng=rng, cells_lo=300, cells_hi=3000, nsteps=3, dt=10.0, seed0=100)
    for r in reqs:
        r.keep = True
    budget = 2 * max(dx.estimate_bytes(2, 3, r.n) for r in reqs)
    srv = serve.Server(hbm_budget=budget, linear=LIN, prep_threads=2)
    peak = []
    srv_done = srv.run(reqs, on_phase=lambda s: peak.append(len(s.batch)))
    srv.close()
    assert len(srv_done) == 6 and all(r.finished == 1 for r in srv_done)
    assert max(peak) >= 2  # requests overlapped on the device
    for r in srv_done:
        prob = r.problem()
        asm = dx.FVAssembler(prob)
        asm.setSolution(prob.u0)
        b = dx.Batch(linear=LIN)
        b.add(asm, nsteps=r.nsteps, dt=r.dt)
        b.run()
        f, info = b.status(asm)
        assert f == 1
        assert np.array_equal(bits(asm.solution()), bits(r.solution)), r.rid
        assert info.newton_steps == r.info.newton_steps
        b.close()
        asm.close()
    st = serve.summarize(srv_done)
    assert st["p99_s"] >= st["p50_s"] > 0


def test_ill_conditioned_problem_falls_back_to_fp64():
    # 1p 2D 200^2 with log10 K ~ N(-11.5, 0.75): cond(D^-1 J) * eps32 is not small, the FP32
    # refinement stalls; the problem continues with FP64 solves and still meets the bar, while a
    # well-conditioned problem in the same batch stays FP32 and bitwise equal to running alone
    kw = dict(seed=5094, mu_log10K=-11.5, sigma_log10K=0.75)
    prob = dx.Problem(kind=1, dim=2, n=(200, 200), **kw)
    asm = dx.FVAssembler(prob)
    asm.setSolution(prob.u0)
    p2, a2 = make("2p3d")
    b = dx.Batch(linear=LIN)
    b.add(asm, nsteps=1, dt=1.0, linear_pde=True, reduction=1e-13)
    b.add(a2, nsteps=SPECS["2p3d"][4], dt=SPECS["2p3d"][5])
    b.run()
    f, info = b.status(asm)
    assert f == 1 and info.fp64_fallback == 1
    f2, info2 = b.status(a2)
    assert f2 == 1 and info2.fp64_fallback == 0
    u = asm.solution()
    u2 = a2.solution()
    b.close()
    asm.close()
    a2.close()
    po = O.OracleProblem(kind=1, dim=2, n=(200, 200), **kw)
    # residual of the solved linear system J du = r(u0) (J the FD Jacobian at u0) at the batched
    # solution, in FP64
    import scipy.sparse as sp
    r0, J0 = po.assemble(po.u0)
    rp, ci = po.pattern()
    Am = sp.csr_matrix((J0, ci, rp), shape=(po.N, po.N))
    res = np.linalg.norm(r0 - Am @ (po.u0 - u)) / np.linalg.norm(r0)
    u_o, info_o = po.newton(po.u0, None, 1.0, linear_pde=1, lin_reduction=1e-14, lin_max_iter=20000)
    assert info_o.converged
    err = np.max(np.abs(u - u_o) / np.abs(u_o))
    print(f"fp64 fallback: true residual {res:.3e}, pw error vs oracle {err:.3e}, "
          f"refinements {info.refinements}, iterations {info.linear_iterations}")
    assert res <= 2e-13
    assert err <= 1e-8
    assert np.array_equal(bits(u2), bits(run_batch(["2p3d"])["2p3d"][0]))


def test_batch_adaptive_time_steps_match_oracle():
    prob, asm = make("2p3d")
    b = dx.Batch(linear=LIN, adaptive_dt=1, max_dt=25.0)
    b.add(asm, nsteps=6, dt=10.0)
    b.run()
    f, info = b.status(asm)
    assert f == 1
    u = asm.solution()
    b.close()
    asm.close()
    kind, dim, n, kw, *_ = SPECS["2p3d"]
    po = O.OracleProblem(kind=kind, dim=dim, n=n, **kw)
    u_o, info_o = po.time_loop(po.u0, 6, 10.0, lin_reduction=1e-10, lin_max_iter=1000, adaptive_dt=1, max_dt=25.0)
    assert info_o.converged
    assert info.time_steps == info_o.time_steps < 6
    assert np.max(np.abs(u[0::2] - u_o[0::2]) / np.abs(u_o[0::2])) <= 1e-8
    assert np.max(np.abs(u[1::2] - u_o[1::2]) / np.maximum(1.0, np.abs(u_o[1::2]))) <= 1e-8


def test_batch_reference_lens_problem():
    # the reference's 2p lens problem (element-wise materials, analytic Jacobian, adaptive dt) as a
    # batch member: bitwise the same alone and next to another problem, and within 1e-8 of the
    # oracle
    from tests.helpers import lens_problem
    po = O.OracleProblem(kind=2, dim=2, n=(48, 32), diff_method=1)
    lens_problem(po)
    u_o, info_o = po.time_loop(po.u0, 12, 250.0, lin_reduction=1e-10, lin_max_iter=1000, adaptive_dt=1)
    assert info_o.converged and info_o.time_steps == 7
    out = []
    for others in ((), ("2p3d",)):
        b = dx.Batch(linear=LIN, adaptive_dt=1)
        pg = dx.Problem(kind=2, dim=2, n=(48, 32), diff_method=dx.DIFF_ANALYTIC)
        lens_problem(pg)
        A = dx.FVAssembler(pg)
        A.setSolution(pg.u0)
        b.add(A, nsteps=12, dt=250.0)
        extra = []
        for nm in others:
            prob, asm = make(nm)
            _, _, _, _, nsteps, dt, _ = SPECS[nm]
            b.add(asm, nsteps=nsteps, dt=dt)
            extra.append(asm)
        b.run()
        f, info = b.status(A)
        assert f == 1 and info.time_steps == 7
        out.append(A.solution())
        for asm in extra:
            assert b.status(asm)[0] == 1
            asm.close()
        b.close()
        A.close()
    assert np.array_equal(bits(out[0]), bits(out[1]))
    u = out[0]
    assert np.max(np.abs(u[0::2] - u_o[0::2]) / np.abs(u_o[0::2])) <= 1e-8
    assert np.max(np.abs(u[1::2] - u_o[1::2]) / np.maximum(1.0, np.abs(u_o[1::2]))) <= 1e-8


def test_launch_shapes_do_not_change_results():
    # the autotuner only picks among launch shapes; every shape gives the same bits
    out = []
    for shape in (None, (1, 1024, 2048), (8, 8192, 16384), "auto"):
        prob, asm = make("2p3d")
        b = dx.Batch(linear=LIN)
        if isinstance(shape, tuple):
            b.set_tuning(*shape)
        b.add(asm, nsteps=2, dt=10.0)
        b.run()
        if shape == "auto":
            c, vg, sg, us = b.autotune(reps=3)
            assert c in (1, 2, 4, 8) and us > 0
            assert b.tuning() == (c, vg, sg)
            asm.setSolution(prob.u0)
            b.remove(asm)
            b.add(asm, nsteps=2, dt=10.0)
            b.run()
        out.append(asm.solution())
        b.close()
        asm.close()
    for u in out[1:]:
        assert np.array_equal(bits(u), bits(out[0]))
