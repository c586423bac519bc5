"""Pins of the oracle's physics and discretization (CPU): reference golden data, closed forms,
invariants and brute-force restatements that a dropped term, a wrong sign or index, or a
transposed operand would fail."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle_ffi as O
from tests.helpers import dp, ensure_oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module", autouse=True)
def _lib():
    ensure_oracle()


def vg_problem(alpha=6.66e-5, n=3.652, l=0.5, swr=0.1, snr=0.1):
    P = O.Problem()
    P.vg_alpha, P.vg_n, P.vg_l, P.swr, P.snr = alpha, n, l, swr, snr
    P.pc_low_swe, P.pc_high_swe, P.krn_low_swe, P.krw_high_swe = 0.01, 0.99, 0.1, 0.9
    return P


def vg_eval(P, sw, reg=True):
    import ctypes as C
    sw = np.ascontiguousarray(sw, dtype=np.float64)
    pc, krw, krn, dpc = (np.empty_like(sw) for _ in range(4))
    if reg:
        O.lib().orc_vg_eval(C.byref(P), len(sw), dp(sw), dp(pc), dp(krw), dp(krn), dp(dpc))
        return pc, krw, krn, dpc
    O.lib().orc_vg_eval_noreg(C.byref(P), len(sw), dp(sw), dp(pc), dp(krw), dp(krn))
    return pc, krw, krn


# ---- van Genuchten (test/material/fluidmatrixinteractions/2p/test_material_2p_vangenuchten.cc) ----

def test_vg_pc_matches_reference_data():
    ref = np.loadtxt(os.path.join(GOLDEN, "test_pcsw_vangenuchten-reference.dat"))
    delta = (1.0 - 0.0) / 99.0                       # dumux/common/math.hh:611-623 linspace
    sw = np.array([0.0 + i * delta for i in range(100)])
    pc = vg_eval(vg_problem(), sw)[0]
    rel = np.abs(pc - ref) / np.abs(ref)
    # the reference values come from glibc pow; the deterministic pow differs by < 1 ulp, amplified
    # at most ~70x by the cancellation in pow(swe, -1/m) - 1 near swe = 0.99
    assert rel.max() < 1e-14, rel.max()
    assert (pc == ref).sum() >= 80


def test_vg_regularized_equals_plain_law_inside_thresholds():
    P = vg_problem()
    lo, hi = 0.01 * (1 - 0.2) + 0.1, 0.99 * (1 - 0.2) + 0.1      # sweToSw of the thresholds
    sw = np.linspace(lo, hi, 200)[1:-1]
    pc_r, krw_r, krn_r, _ = vg_eval(P, sw)
    pc_n, krw_n, krn_n = vg_eval(P, sw, reg=False)
    assert np.array_equal(pc_r, pc_n)
    swe = (sw - 0.1) / 0.8
    inside_kr = (swe < 0.9) & (swe > 0.1)
    assert np.array_equal(krw_r[inside_kr], krw_n[inside_kr])
    assert np.array_equal(krn_r[inside_kr], krn_n[inside_kr])


def test_vg_regularization_end_conditions_and_continuity():
    P = vg_problem(alpha=4e-4, n=2.5, swr=0.05, snr=0.0)
    s = lambda swe: swe * (1 - 0.05) + 0.05
    pc, krw, krn, dpc = vg_eval(P, np.array([s(1.0), s(1.2), s(0.0), s(-0.1)]))
    assert pc[0] == 0.0 and pc[1] < 0.0                 # linear continuation beyond swe = 1
    assert krw[0] == 1.0 and krw[1] == 1.0 and krn[0] == 0.0
    assert krn[2] == 1.0 and krn[3] == 1.0 and krw[2] == 0.0
    # C0 / C1 continuity across every threshold (spline end slopes equal the law's slopes)
    for swe_t in (0.01, 0.99, 0.1, 0.9):
        e = 1e-9
        a = vg_eval(P, np.array([s(swe_t - e), s(swe_t + e)]))
        for f in a[:3]:
            assert abs(f[1] - f[0]) <= 1e-6 * max(1.0, abs(f[0])), (swe_t, f)


def test_vg_dpc_dsw_matches_central_differences():
    # testDerivatives of testmateriallawfunctions.hh:24-43: central differences, 10 % tolerance
    P = vg_problem()
    sw = np.linspace(0.0, 1.0, 100)
    h = 1e-8
    pcp = vg_eval(P, sw + h)[0]
    pcm = vg_eval(P, sw - h)[0]
    dpc = vg_eval(P, sw)[3]
    num = (pcp - pcm) / (2 * h)
    ok = np.abs(dpc - num) <= 0.1 * np.maximum(np.abs(dpc), np.abs(num))
    swe = (sw - 0.1) / 0.8
    kinks = np.isclose(swe, 0.01, atol=2e-8) | np.isclose(swe, 0.99, atol=2e-8) | np.isclose(swe, 1.0, atol=2e-8)
    assert ok[~kinks].all()


def vg_deriv(P, sw):
    import ctypes as C
    sw = np.ascontiguousarray(sw, dtype=np.float64)
    out = [np.empty_like(sw) for _ in range(3)]
    O.lib().orc_vg_eval_deriv(C.byref(P), len(sw), dp(sw), *[dp(o) for o in out])
    return out


def test_vg_regularized_kr_derivatives():
    # dkrw/dsw and dkrn/dsw (materiallaw.hh:195-240, vangenuchten.hh:523-563) against central
    # differences of krw / krn on every branch: base law, the splines above krwHighSwe = 0.9 and
    # below krnLowSwe = 0.1, and the constant continuations outside [0, 1], where they vanish
    P = vg_problem(alpha=4e-4, n=2.5, swr=0.05, snr=0.0)
    s = lambda swe: swe * (1 - 0.05) + 0.05
    swe = np.linspace(-0.05, 1.05, 443)
    sw = s(swe)
    dpc, dkrw, dkrn = vg_deriv(P, sw)
    assert np.array_equal(dpc, vg_eval(P, sw)[3])
    h = 1e-7
    kp, km = vg_eval(P, sw + h), vg_eval(P, sw - h)
    near = lambda t: np.abs(swe - t) < 1e-5
    kinks = near(0.0) | near(0.1) | near(0.9) | near(1.0)
    for k, d in ((1, dkrw), (2, dkrn)):
        num = (kp[k] - km[k]) / (2 * h)
        assert np.allclose(d[~kinks], num[~kinks], rtol=2e-5, atol=1e-6), k
    out = (swe <= 0.0) | (swe >= 1.0)
    assert np.all(dkrw[out] == 0.0) and np.all(dkrn[out] == 0.0)
    assert np.all(dkrw[~out] >= 0.0) and np.all(dkrn[~out] <= 0.0)
    # C1 at the spline joints: the one-sided derivatives agree
    for t in (0.1, 0.9):
        a, b = vg_deriv(P, np.array([s(t - 1e-9), s(t + 1e-9)]))[1:]
        assert abs(a[1] - a[0]) <= 1e-6 * abs(a[0]) and abs(b[1] - b[0]) <= 1e-6 * abs(b[0]), t


# ---- pattern / coloring ------------------------------------------------------------------

@pytest.mark.parametrize("dim,n", [(2, (7, 5)), (3, (4, 5, 3)), (3, (1, 6, 4)), (2, (1, 9))])
def test_pattern_brute_force(dim, n):
    p = O.OracleProblem(kind=1, dim=dim, n=n)
    rp, ci = p.pattern()
    nn = list(n) + [1] * (3 - dim)
    idx = np.array(np.unravel_index(np.arange(p.N), nn[::-1]))[::-1]  # x-fastest (i, j, k)
    for I in range(p.N):
        cols = {I}
        for a in range(dim):
            for s in (-1, 1):
                q = idx[:, I].copy()
                q[a] += s
                if 0 <= q[a] < nn[a]:
                    cols.add(int(q[0] + nn[0] * (q[1] + nn[1] * q[2])))
        assert list(ci[rp[I]:rp[I + 1]]) == sorted(cols)


def test_coloring_is_distance_two():
    p = O.OracleProblem(kind=1, dim=3, n=(6, 5, 4))
    nc, colors = p.coloring()
    rp, ci = p.pattern()
    assert colors.min() == 0 and colors.max() == nc - 1
    for I in range(p.N):
        nb = [j for j in ci[rp[I]:rp[I + 1]] if j != I]
        reach = set(nb)
        for j in nb:
            reach.update(int(k) for k in ci[rp[j]:rp[j + 1]])
        reach.discard(I)
        assert all(colors[k] != colors[I] for k in reach)


# ---- closed forms -----------------------------------------------------------------------

def onep(n, K, gravity=0, p_left=2e5, p_right=1e5):
    p = O.OracleProblem(kind=1, dim=2, n=n, gravity=gravity, p_left=p_left, p_right=p_right)
    p.set_arrays(K=K)
    # linear problem: a large difference step removes the forward-difference noise, as the
    # reference's own 1p test does (test/porousmediumflow/1p/incompressible/params.input)
    p.P.fd_base_eps = 0.1
    p.P.fd_magnitude[0] = 1e5
    return p


def solve_linear(p, u0=None):
    u, info = p.newton(p.u0 if u0 is None else u0, None, 1.0, linear_pde=1, lin_reduction=1e-14,
                       lin_max_iter=5000)
    assert info.converged
    return u


def centers(n):
    h = 1.0 / n
    return np.array([0.0 + i * h + 0.5 * h for i in range(n)])


def test_onep_homogeneous_linear_pressure():
    n = (16, 8)
    p = onep(n, np.full(128, 3e-11))
    u = solve_linear(p).reshape(n[1], n[0])
    x = centers(n[0])
    exact = 2e5 + (1e5 - 2e5) * x
    assert np.max(np.abs(u - exact[None, :]) / exact[None, :]) < 1e-11


def test_onep_series_resistance():
    # K varying along x only: discrete solution = harmonic series of half-cell resistances
    n = (12, 3)
    rng = np.random.default_rng(3)
    kx = 10 ** rng.normal(-11, 1, n[0])
    p = onep(n, np.tile(kx, n[1]))
    u = solve_linear(p).reshape(n[1], n[0])
    h = 1.0 / n[0]
    R = h / kx                                   # each cell: two half-cell resistances h/2/K
    total = R.sum()
    q = (2e5 - 1e5) / total
    exact = 2e5 - q * (np.cumsum(R) - 0.5 * R)
    assert np.max(np.abs(u - exact[None, :]) / exact[None, :]) < 1e-11


def textbook_onep_matrix(p, n, K, rho=1000.0, mu=1e-3, gravity=True):
    """TPFA matrix and right-hand side of the OneP problem from first principles."""
    nx, ny = n
    h = (1.0 / nx, 1.0 / ny)
    lam = rho / mu
    g = 9.81 if gravity else 0.0
    A = sp.lil_matrix((nx * ny, nx * ny))
    b = np.zeros(nx * ny)
    for j in range(ny):
        for i in range(nx):
            I = i + nx * j
            for a, s in ((0, -1), (0, 1), (1, -1), (1, 1)):
                q = [i, j]
                q[a] += s
                area = h[1 - a]
                if 0 <= q[a] < n[a]:
                    J = q[0] + nx * q[1]
                    t = area / (0.5 * h[a] / K[I] + 0.5 * h[a] / K[J])
                    A[I, I] += t * lam
                    A[I, J] -= t * lam
                    if a == 1:  # gravity drives flux through horizontal faces
                        b[I] -= lam * t * (-s) * rho * g * h[1] * (-1)
                elif a == 0:   # Dirichlet left / right at the face
                    t = area * K[I] / (0.5 * h[a])
                    A[I, I] += t * lam
                    b[I] += t * lam * (2e5 if s < 0 else 1e5)
    return A.tocsr(), b


def test_onep_fd_jacobian_equals_textbook_tpfa_matrix():
    n = (9, 7)
    rng = np.random.default_rng(5)
    K = 10 ** rng.normal(-11, 0.8, n[0] * n[1])
    p = onep(n, K, gravity=0)
    r, J = p.assemble(p.u0)
    rp, ci = p.pattern()
    Afd = sp.csr_matrix((J, ci, rp), shape=(p.N, p.N))
    Atb, b = textbook_onep_matrix(p, n, K, gravity=False)
    diff = abs(Afd - Atb).max()
    assert diff <= 1e-6 * abs(Atb).max(), diff
    # symmetric, zero row sums away from the Dirichlet columns
    assert abs(Afd - Afd.T).max() <= 1e-6 * abs(Atb).max()


def test_onep_analytic_jacobian_equals_textbook_tpfa_matrix():
    # CCLocalAssembler<analytic> with OnePIncompressibleLocalResidual: d r_I / d p_J = -tij rho/mu,
    # d r_I / d p_I = sum of tij rho/mu over the faces incl. Dirichlet ones: the textbook matrix,
    # up to rounding (no difference quotient involved)
    n = (9, 7)
    rng = np.random.default_rng(5)
    K = 10 ** rng.normal(-11, 0.8, n[0] * n[1])
    p = onep(n, K, gravity=1)
    r, J = p.assemble(p.u0, diff_method=1)
    assert np.array_equal(r, p.assemble(p.u0)[0])
    rp, ci = p.pattern()
    A = sp.csr_matrix((J, ci, rp), shape=(p.N, p.N))
    Atb, _ = textbook_onep_matrix(p, n, K, gravity=True)
    assert abs(A - Atb).max() <= 1e-13 * abs(Atb).max()
    assert abs(A - A.T).max() <= 1e-13 * abs(Atb).max()


ANALYTIC_CASES = [(2, 3, (10, 9, 8), dict(seed=3)), (2, 2, (20, 16), dict(seed=4)),
                  (2, 3, (8, 8, 8), dict(seed=5, gravity=0)), (1, 2, (20, 20), dict(seed=1)),
                  (1, 3, (6, 5, 4), dict(seed=2))]


@pytest.mark.parametrize("kind,dim,n,kw", ANALYTIC_CASES)
def test_analytic_jacobian_matches_forward_differences(kind, dim, n, kw):
    # the hand-coded derivatives (1p/2p incompressiblelocalresidual.hh) are the exact derivatives
    # of the residual away from upwind switches: they agree with the numeric Jacobian to its
    # difference-quotient noise (~1e-6 of each column's scale). A dropped term (storage, the
    # capillary-pressure or relative-permeability part, an outside weight) or a wrong sign shows
    # as an O(1) difference. The residual is the same function in both assemblers.
    from tests.helpers import perturbed_state
    po = O.OracleProblem(kind=kind, dim=dim, n=n, **kw)
    u = perturbed_state(po, po.u0, seed=3, sn_max=0.3)
    m = po.m
    for uprev in (None, po.u0):
        r_n, J_n = po.assemble(u, uprev, 10.0, diff_method=0)
        r_a, J_a = po.assemble(u, uprev, 10.0, diff_method=1)
        assert np.array_equal(r_n.view(np.uint64), r_a.view(np.uint64))
        Jn, Ja = J_n.reshape(-1, m, m), J_a.reshape(-1, m, m)
        scale = np.abs(Jn).max(axis=(0, 1))
        err = (np.abs(Ja - Jn) / scale[None, None, :]).max(axis=(0, 1))
        assert np.all(err <= 1e-5), err


def test_analytic_newton_time_loop_matches_numeric():
    p = O.OracleProblem(kind=2, dim=3, n=(8, 8, 6), seed=21)
    pa = O.OracleProblem(kind=2, dim=3, n=(8, 8, 6), seed=21, diff_method=1)
    u_n, i_n = p.time_loop(p.u0, 3, 10.0, lin_reduction=1e-12, lin_max_iter=2000)
    u_a, i_a = pa.time_loop(pa.u0, 3, 10.0, lin_reduction=1e-12, lin_max_iter=2000)
    assert i_n.converged and i_a.converged
    assert abs(i_a.newton_steps - i_n.newton_steps) <= 1
    assert np.max(np.abs(u_a[0::2] - u_n[0::2]) / np.abs(u_n[0::2])) <= 1e-8
    assert np.max(np.abs(u_a[1::2] - u_n[1::2])) <= 1e-8
    assert u_n[1::2].max() > 1e-3


def test_onep_solution_matches_spsolve_of_textbook_system_with_gravity():
    n = (10, 6)
    rng = np.random.default_rng(8)
    K = 10 ** rng.normal(-11, 0.8, n[0] * n[1])
    p = onep(n, K, gravity=1)
    u = solve_linear(p)
    # Newton on a linear problem: J du = r(u0) -> u = u0 - du; the textbook system includes gravity
    r, J = p.assemble(p.u0)
    rp, ci = p.pattern()
    x = spla.spsolve(sp.csr_matrix((J, ci, rp), shape=(p.N, p.N)).tocsc(), r)
    assert np.max(np.abs(u - (p.u0 - x)) / np.abs(u)) < 1e-10
    r_after, _ = p.assemble(u, jacobian=False)
    assert np.max(np.abs(r_after)) < 1e-10 * np.max(np.abs(r))


@pytest.mark.parametrize("diff_method", [0, 1])
def test_onep_reference_lens_problem_matches_reference_vtu(diff_method):
    # test_1p_incompressible_tpfa (analytic Jacobian, test/porousmediumflow/1p/incompressible/
    # CMakeLists.txt:19-27) and its numeric-difference variant share the reference VTU
    ref = np.loadtxt(os.path.join(GOLDEN, "test_1p_cc_reference_p.txt"))
    n = (10, 10)
    p = O.OracleProblem(kind=1, dim=2, n=n, gravity=1)
    c = centers(10)
    eps = 1.5e-7
    inlens = lambda v: (v >= 0.2 + eps) & (v <= 0.8 - eps)
    K = np.where(inlens(c)[None, :] & inlens(c)[:, None], 1e-12, 1e-10).ravel()
    bct = np.zeros(p.nbf, dtype=np.int32)
    bcv = np.zeros(p.nbf)
    # sides -x (0..9), +x (10..19), -y (20..29), +y (30..39): Dirichlet on -y and +y
    bct[20:40] = 1
    bcv[20:30] = 1e5 + (-1e5) * (0.0 - 1.0)
    bcv[30:40] = 1e5 + (-1e5) * (1.0 - 1.0)
    p.set_arrays(K=K, bc_type=bct, bc_value=bcv)
    p.P.fd_base_eps = 0.1                      # Assembly.NumericDifference.BaseEpsilon
    p.P.fd_magnitude[0] = 1e5                  # Assembly.NumericDifference.PriVarMagnitude
    p.P.diff_method = diff_method
    u, info = p.newton(np.zeros(p.N), None, 1.0, linear_pde=1, lin_reduction=1e-14, lin_max_iter=5000)
    assert info.converged
    assert np.max(np.abs(u - ref)) <= 1.0, np.max(np.abs(u - ref))   # 6-digit Float32 ascii print


def lens_pc(p, sw, mat):
    """Capillary pressure of each cell with its material (0 outside, 1 lens)."""
    pc = np.empty_like(sw)
    for mm, kw in ((0, dict(alpha=0.0037, n=4.7, swr=0.05, snr=0.0)), (1, dict(alpha=0.00045, n=7.3, swr=0.18, snr=0.0))):
        sel = mat == mm
        pc[sel] = vg_eval(vg_problem(**kw), sw[sel])[0]
    return pc


@pytest.mark.parametrize("diff_method", [1, 0])
def test_twop_reference_lens_problem_matches_reference_vtu(diff_method):
    # test_2p_incompressible_tpfa_analytic / _tpfa (test/porousmediumflow/2p/incompressible/
    # CMakeLists.txt:5-27): 48x32 lens problem, two van Genuchten materials, gravity, Dirichlet
    # and Neumann sides, adaptive dt (suggestTimeStepSize) from 250 s to t_end 3000 s, Newton's
    # linear tolerance 1e-6 (newtonsolver.hh:232). The 7th output must pass the reference's own
    # fuzzy comparison; the agreement is in fact ~5e-6 relative (the VTU prints Float32).
    from tests.helpers import fuzzy_close, lens_problem, lens_reference
    p = O.OracleProblem(kind=2, dim=2, n=(48, 32), diff_method=diff_method)
    mat = lens_problem(p)
    u, info = p.time_loop(p.u0, 12, 250.0, lin_reduction=1e-6, lin_max_iter=500, adaptive_dt=1)
    assert info.converged and info.time_steps == 7
    ref = lens_reference()
    pw, sn = u[0::2], u[1::2]
    pc = lens_pc(p, 1 - sn, mat)
    mine = {"S_aq": 1 - sn, "p_aq": pw, "S_napl": sn, "p_napl": pw + pc, "pc": pc}
    for nm, a in mine.items():
        assert fuzzy_close(a, ref[nm]), nm
    assert np.max(np.abs(pw - ref["p_aq"]) / ref["p_aq"]) < 1e-5
    assert np.max(np.abs(sn - ref["S_napl"])) < 1e-5
    assert sn.max() > 0.3                    # the DNAPL has spread into the domain


# ---- 2p: Jacobian consistency and conservation ------------------------------------------

def test_twop_jacobian_matches_central_differences():
    p = O.OracleProblem(kind=2, dim=2, n=(6, 5), seed=3)
    rng = np.random.default_rng(2)
    u = p.u0.copy()
    u[1::2] = rng.uniform(0.15, 0.6, p.N)     # away from the regularization kinks
    # pressures off hydrostatic equilibrium so that no face flux sits at an upwind switch
    u[0::2] += rng.uniform(-500.0, 500.0, p.N)
    uprev = p.u0.copy()
    r0, J = p.assemble(u, uprev, 10.0)
    rp, ci = p.pattern()
    Jd = sp.bsr_matrix((J.reshape(-1, 2, 2), ci, rp), shape=(2 * p.N, 2 * p.N)).toarray()
    for col in rng.choice(2 * p.N, 12, replace=False):
        h = 1e-3 if col % 2 == 0 else 1e-6
        up, um = u.copy(), u.copy()
        up[col] += h
        um[col] -= h
        cd = (p.assemble(up, uprev, 10.0, jacobian=False)[0] - p.assemble(um, uprev, 10.0, jacobian=False)[0]) / (2 * h)
        assert np.allclose(Jd[:, col], cd, rtol=2e-3, atol=1e-6 * np.abs(cd).max()), col


def test_twop_mass_balance_after_converged_step():
    p = O.OracleProblem(kind=2, dim=3, n=(8, 8, 8), seed=4)
    u, info = p.time_loop(p.u0, 1, 10.0, lin_reduction=1e-12, max_rel_shift=1e-10)
    assert info.converged
    r, _ = p.assemble(u, p.u0, 10.0, jacobian=False)
    # the converged residual is tiny against the storage change it balances
    V = (1.0 / 8) ** 3
    storage = 0.2 * 1460.0 * u[1::2] * V / 10.0
    assert np.abs(r[1::2]).sum() < 1e-9 * storage.sum()
    # DNAPL mass gained = inflow through the inlet faces (the residual sums telescope)
    inlet_area = (4 * 4) * (1.0 / 8) ** 2
    assert abs(storage.sum() - 0.04 * inlet_area) < 1e-8 * 0.04 * inlet_area


# ---- linear algebra ---------------------------------------------------------------------

def small_system(seed=1):
    p = O.OracleProblem(kind=2, dim=3, n=(7, 6, 5), seed=seed)
    rng = np.random.default_rng(seed)
    u = p.u0.copy()
    u[1::2] = rng.uniform(0.0, 0.5, p.N)
    r, J = p.assemble(u, p.u0, 10.0)
    rp, ci = p.pattern()
    return p, r, J, rp, ci


def test_ilu0_defining_property():
    p, r, J, rp, ci = small_system()
    lu = O.ilu0(rp, ci, J, 2)
    m = 2
    B = lambda k: lu[k * 4:(k + 1) * 4].reshape(2, 2)
    Ablk = {(i, ci[k]): J[k * 4:(k + 1) * 4].reshape(2, 2) for i in range(p.N) for k in range(rp[i], rp[i + 1])}
    L = {(i, ci[k]): B(k) for i in range(p.N) for k in range(rp[i], rp[i + 1]) if ci[k] < i}
    U = {(i, ci[k]): B(k) for i in range(p.N) for k in range(rp[i], rp[i + 1]) if ci[k] > i}
    D = {i: np.linalg.inv(B(k)) for i in range(p.N) for k in range(rp[i], rp[i + 1]) if ci[k] == i}
    # (L + I) (D + U) restricted to the pattern equals A
    for (i, j), a in Ablk.items():
        s = np.zeros((m, m))
        for k in range(min(i, j) + 1):
            lik = np.eye(m) if k == i else L.get((i, k))
            if lik is None:
                continue
            ukj = D[k] if k == j else U.get((k, j))
            if ukj is None:
                continue
            s += lik @ ukj
        assert np.allclose(s, a, rtol=1e-10, atol=1e-12 * np.abs(a).max() + 1e-300), (i, j)


def test_ilu0_apply_solves_factor_system():
    p, r, J, rp, ci = small_system(2)
    lu = O.ilu0(rp, ci, J, 2)
    v = O.ilu0_apply(rp, ci, lu, r, 2)
    # rebuild (L+I)(D+U) as a matrix and check (L+I)(D+U) v = r
    n = p.N
    Lm = sp.lil_matrix((2 * n, 2 * n))
    Um = sp.lil_matrix((2 * n, 2 * n))
    for i in range(n):
        Lm[2 * i:2 * i + 2, 2 * i:2 * i + 2] = np.eye(2)
        for k in range(rp[i], rp[i + 1]):
            j = ci[k]
            blk = lu[4 * k:4 * k + 4].reshape(2, 2)
            if j < i:
                Lm[2 * i:2 * i + 2, 2 * j:2 * j + 2] = blk
            elif j > i:
                Um[2 * i:2 * i + 2, 2 * j:2 * j + 2] = blk
            else:
                Um[2 * i:2 * i + 2, 2 * i:2 * i + 2] = np.linalg.inv(blk)
    lhs = Lm.tocsr() @ (Um.tocsr() @ v)
    assert np.allclose(lhs, r, rtol=1e-9, atol=1e-12 * np.abs(r).max())


@pytest.mark.parametrize("precond", [1, 2])
def test_bicgstab_matches_spsolve(precond):
    p, r, J, rp, ci = small_system(3)
    x, info = O.bicgstab(rp, ci, J, r, 2, reduction=1e-12, maxit=2000, precond=precond)
    assert info.converged
    A = sp.bsr_matrix((J.reshape(-1, 2, 2), ci, rp), shape=(2 * p.N, 2 * p.N)).tocsc()
    xs = spla.spsolve(A, r)
    assert np.max(np.abs(x - xs)) / np.max(np.abs(xs)) < 1e-8
    assert np.linalg.norm(r - A @ x) / np.linalg.norm(r) < 1e-12


def test_bicgstab_reports_non_convergence():
    p, r, J, rp, ci = small_system(4)
    x, info = O.bicgstab(rp, ci, J, r, 2, reduction=1e-14, maxit=2, precond=2)
    assert not info.converged


def test_newton_time_step_retry_on_linear_failure():
    # LinearSolver.MaxIterations 1 forces NumericalProblem -> dt halving until the retry limit
    p = O.OracleProblem(kind=2, dim=2, n=(6, 6), seed=5)
    u, info = p.time_loop(p.u0, 1, 10.0, lin_reduction=1e-12, lin_max_iter=1, max_dt_divisions=2)
    assert not info.converged and info.dt_retries == 2


def test_suggest_time_step_special_cases():
    # suggestTimeStepSize (newtonsolver.hh:784-798) at the points its description fixes: on
    # target keep dt; k*target iterations divide dt by k; zero iterations grow dt by 1/1.2;
    # strictly decreasing in the iteration count
    s = O.lib().orc_suggest_time_step
    assert s(7.0, 10, 10) == 7.0
    assert s(8.0, 20, 10) == 4.0 and s(9.0, 30, 10) == 3.0
    assert abs(s(6.0, 0, 10) - 6.0 * (1.0 + 1.0 / 1.2)) <= 1e-15
    vals = [s(1.0, k, 10) for k in range(0, 40)]
    assert all(a > b for a, b in zip(vals, vals[1:]))


def test_adaptive_time_loop_grows_dt_and_ends_on_time():
    # with few Newton iterations per step (well below TargetSteps) the adaptive loop takes fewer
    # steps than the fixed one, never overshoots the end time, and respects max_dt
    p = O.OracleProblem(kind=2, dim=2, n=(8, 8), seed=3)
    _, fixed = p.time_loop(p.u0, 6, 10.0, lin_reduction=1e-10, lin_max_iter=500)
    _, adapt = p.time_loop(p.u0, 6, 10.0, lin_reduction=1e-10, lin_max_iter=500, adaptive_dt=1)
    _, capped = p.time_loop(p.u0, 6, 10.0, lin_reduction=1e-10, lin_max_iter=500, adaptive_dt=1, max_dt=12.0)
    assert fixed.converged and adapt.converged and capped.converged
    assert fixed.time_steps == 6
    assert adapt.time_steps < fixed.time_steps
    assert capped.time_steps >= 5  # 60 s in steps of at most 12 s
