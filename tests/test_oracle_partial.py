"""Partial reassembly in the oracle (PartialReassembler for CCTpfa, dumux/assembly/
partialreassembler.hh:301-375, with NewtonSolver's hooks, dumux/nonlinear/newtonsolver.hh:415-419,
565-603, 1241-1244): known answers for the colors, the column bookkeeping of the assembly, and
the reference's own 2p lens test, which runs with Newton.EnablePartialReassembly = true
(test/porousmediumflow/2p/incompressible/params.input)."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle_ffi as O
from tests.helpers import ensure_oracle, fuzzy_close, lens_problem, lens_reference, perturbed_state


@pytest.fixture(scope="module", autouse=True)
def _lib():
    ensure_oracle()


def test_colors_known_answer():
    # three cells, m = 2, values exact in binary; relative shift |a-b| / max(1, |a+b|/2) per
    # unknown, the cell's shift is the max over its unknowns
    u_last = np.array([8.0, 0.0, 0.0, 0.5, 2.0, 1.0])
    u = np.array([12.0, 0.0, 0.25, 0.5, 2.0, 1.0])
    dist = np.array([0.0, 0.25, 0.125])
    green, d, ng = O.partial_colors(u_last, u, 2, 0.5, dist)
    # cell 0: |8-12| / 10 = 0.4, not above 0.5 -> green, distance 0.4
    # cell 1: 0.25 + |0-0.25| / 1 = 0.5, exactly the threshold, not above -> green
    # cell 2: no change, 0.125 -> green
    assert green.tolist() == [1, 1, 1] and ng == 3
    assert d.tolist() == [0.4, 0.5, 0.125]
    green, d, ng = O.partial_colors(u_last, u, 2, 0.45, dist)
    # threshold 0.45: cell 1 (0.5) turns red and its distance is reset
    assert green.tolist() == [1, 0, 1] and ng == 2
    assert d.tolist() == [0.4, 0.0, 0.125]


def test_colors_use_the_reference_update_arithmetic():
    # the shift is taken between uLast and uLast - (uLast - u), as newtonUpdate forms it
    # (newtonsolver.hh:593-594, 1270-1276); with operands of different magnitudes the two
    # subtractions round, and a restatement from u directly gives other bits for some cells
    rng = np.random.default_rng(5)
    u_last = rng.uniform(-1, 1, 4000) * 10.0 ** rng.integers(-3, 6, 4000)
    u = rng.uniform(-1, 1, 4000) * 10.0 ** rng.integers(-3, 6, 4000)
    green, d, _ = O.partial_colors(u_last, u, 1, 1e300, np.zeros(4000))
    nxt = u_last - (u_last - u)
    want = np.abs(u_last - nxt) / np.maximum(1.0, np.abs(u_last + nxt) * 0.5)
    direct = np.abs(u_last - u) / np.maximum(1.0, np.abs(u_last + u) * 0.5)
    assert np.sum(want != direct) > 10          # the data tells the two forms apart
    assert np.array_equal(d, want) and green.all()


def _jac(p, J):
    rp, ci = p.pattern()
    m = p.m
    return sp.bsr_matrix((J.reshape(-1, m, m), ci, rp), shape=(m * p.N, m * p.N)).toarray()


def test_assembly_keeps_green_columns_and_recomputes_red_ones():
    p = O.OracleProblem(kind=2, dim=3, n=(6, 5, 4), seed=7)
    u0 = perturbed_state(p, p.u0, seed=1)
    u1 = perturbed_state(p, p.u0, seed=2)
    _, J0 = p.assemble(u0, p.u0, 10.0)
    r1, J1 = p.assemble(u1, p.u0, 10.0)
    green = (np.random.default_rng(3).uniform(size=p.N) < 0.5).astype(np.uint8)
    # all red: the full assembly, bit for bit
    r, J = p.assemble_partial(u1, p.u0, 10.0, np.zeros(p.N, np.uint8), J0)
    assert np.array_equal(r, r1) and np.array_equal(J, J1)
    # all green: the old Jacobian, the new residual
    r, J = p.assemble_partial(u1, p.u0, 10.0, np.ones(p.N, np.uint8), J0)
    assert np.array_equal(r, r1) and np.array_equal(J, J0)
    # mixed: green cells' columns from J0, red cells' columns from J1
    r, J = p.assemble_partial(u1, p.u0, 10.0, green, J0)
    A, A0, A1 = _jac(p, J), _jac(p, J0), _jac(p, J1)
    col_green = np.repeat(green.astype(bool), 2)
    assert np.array_equal(A[:, col_green], A0[:, col_green])
    assert np.array_equal(A[:, ~col_green], A1[:, ~col_green])
    assert np.array_equal(r, r1)


def test_analytic_jacobian_is_rejected():
    # the column reset presumes column-wise (numeric) assembly; the reference's analytic lens
    # test runs with -Newton.EnablePartialReassembly false (CMakeLists.txt:19-27)
    p = O.OracleProblem(kind=2, dim=2, n=(5, 4), diff_method=1)
    with pytest.raises(ValueError):
        p.time_loop(p.u0, 1, 10.0, partial_reassembly=1)


def test_newton_with_partial_reassembly_converges_to_the_full_newton_solution():
    p = O.OracleProblem(kind=2, dim=3, n=(10, 9, 8), seed=11)
    kw = dict(lin_reduction=1e-12, lin_max_iter=2000)
    u_full, i_full = p.time_loop(p.u0, 3, 20.0, **kw)
    u_part, i_part = p.time_loop(p.u0, 3, 20.0, partial_reassembly=1, **kw)
    assert i_full.converged and i_part.converged
    assert i_full.reassembled_cells == i_full.newton_steps * p.N
    assert 0 < i_part.reassembled_cells < i_part.newton_steps * p.N
    assert np.max(np.abs(u_part[0::2] - u_full[0::2]) / np.abs(u_full[0::2])) < 1e-8
    assert np.max(np.abs(u_part[1::2] - u_full[1::2])) < 1e-8
    # the modified Newton extreme (every cell green after the first iteration) still converges
    u_frozen, i_frozen = p.time_loop(p.u0, 3, 20.0, partial_reassembly=1, reassembly_min_threshold=1e300,
                                     reassembly_max_threshold=1e300, **kw)
    assert i_frozen.converged
    assert i_frozen.reassembled_cells <= 3 * p.N + i_frozen.dt_retries * p.N
    assert np.max(np.abs(u_frozen[1::2] - u_full[1::2])) < 1e-7


def test_reference_lens_test_with_partial_reassembly_matches_reference_vtu():
    # test_2p_incompressible_tpfa as configured (params.input: Newton.EnablePartialReassembly =
    # true, numeric Jacobian): the 7th output passes the reference's fuzzy comparison
    p = O.OracleProblem(kind=2, dim=2, n=(48, 32), diff_method=0)
    lens_problem(p)
    u, info = p.time_loop(p.u0, 12, 250.0, lin_reduction=1e-6, lin_max_iter=500, adaptive_dt=1, partial_reassembly=1)
    assert info.converged and info.time_steps == 7
    assert info.reassembled_cells < info.newton_steps * p.N     # columns were kept
    ref = lens_reference()
    assert fuzzy_close(u[0::2], ref["p_aq"]) and fuzzy_close(u[1::2], ref["S_napl"])
    assert np.max(np.abs(u[0::2] - ref["p_aq"]) / ref["p_aq"]) < 1e-5
    assert np.max(np.abs(u[1::2] - ref["S_napl"])) < 1e-5
