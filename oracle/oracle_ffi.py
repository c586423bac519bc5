"""Oracle (test infrastructure only): ctypes binding of oracle/build/liboracle.so.

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, where the
oracle is the checker. The dumux_amd product never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")


class VGMaterial(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("vg_alpha", "vg_n", "vg_l", "swr", "snr", "pc_low_swe", "pc_high_swe",
                                          "krn_low_swe", "krw_high_swe")]


class Problem(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("model", C.c_int32), ("n", C.c_int32 * 3), ("gravity", C.c_int32),
        ("lower", C.c_double * 3), ("upper", C.c_double * 3),
        ("rho", C.c_double * 2), ("mu", C.c_double * 2), ("porosity", C.c_double),
        ("vg_alpha", C.c_double), ("vg_n", C.c_double), ("vg_l", C.c_double),
        ("swr", C.c_double), ("snr", C.c_double),
        ("pc_low_swe", C.c_double), ("pc_high_swe", C.c_double),
        ("krn_low_swe", C.c_double), ("krw_high_swe", C.c_double),
        ("K", C.POINTER(C.c_double)), ("bc_type", C.POINTER(C.c_int32)),
        ("bc_value", C.POINTER(C.c_double)),
        ("fd_base_eps", C.c_double), ("fd_magnitude", C.c_double * 2),
        ("diff_method", C.c_int32),
        ("material_id", C.POINTER(C.c_int32)), ("material1", VGMaterial),
    ]


class BuildSpec(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("dim", C.c_int32), ("n", C.c_int32 * 3),
        ("seed", C.c_uint32), ("pid", C.c_uint32),
        ("mu_log10K", C.c_double), ("sigma_log10K", C.c_double),
        ("gravity", C.c_int32),
        ("p_left", C.c_double), ("p_right", C.c_double), ("p_ref", C.c_double),
        ("influx", C.c_double), ("inlet_lo", C.c_double), ("inlet_hi", C.c_double),
    ]


class SolveInfo(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iterations", C.c_int32),
                ("reduction", C.c_double), ("def0", C.c_double), ("def_", C.c_double)]


class NewtonParams(C.Structure):
    _fields_ = [
        ("max_rel_shift", C.c_double), ("min_steps", C.c_int32), ("max_steps", C.c_int32),
        ("lin_reduction", C.c_double), ("lin_max_iter", C.c_int32), ("precond", C.c_int32),
        ("max_dt_divisions", C.c_int32), ("retry_factor", C.c_double),
        ("linear_pde", C.c_int32), ("nthreads", C.c_int32), ("verbosity", C.c_int32),
        ("adaptive_dt", C.c_int32), ("target_steps", C.c_int32), ("max_dt", C.c_double),
        ("partial_reassembly", C.c_int32), ("reassembly_min_threshold", C.c_double),
        ("reassembly_max_threshold", C.c_double), ("reassembly_shift_weight", C.c_double),
    ]


class NewtonInfo(C.Structure):
    _fields_ = [
        ("converged", C.c_int32), ("newton_steps", C.c_int32), ("wasted_steps", C.c_int32),
        ("linear_iterations", C.c_int32), ("time_steps", C.c_int32), ("dt_retries", C.c_int32),
        ("last_shift", C.c_double), ("t_assemble", C.c_double), ("t_solve", C.c_double),
        ("t_update", C.c_double), ("reassembled_cells", C.c_int64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        dp, ip, i64 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_int64
        L.orc_dpow.restype = C.c_double
        L.orc_dpow.argtypes = [C.c_double, C.c_double]
        L.orc_dlog.restype = C.c_double
        L.orc_dlog.argtypes = [C.c_double]
        L.orc_suggest_time_step.restype = C.c_double
        L.orc_suggest_time_step.argtypes = [C.c_double, C.c_int32, C.c_int32]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_lognormal_K.argtypes = [C.c_uint32, C.c_uint32, i64, C.c_double, C.c_double, dp]
        L.orc_material_params.argtypes = [C.POINTER(Problem), dp]
        L.orc_vg_eval.argtypes = [C.POINTER(Problem), i64, dp, dp, dp, dp, dp]
        L.orc_vg_eval_noreg.argtypes = [C.POINTER(Problem), i64, dp, dp, dp, dp]
        L.orc_num_cells.restype = i64
        L.orc_num_cells.argtypes = [C.POINTER(Problem)]
        L.orc_num_boundary_faces.restype = i64
        L.orc_num_boundary_faces.argtypes = [C.c_int32, ip]
        L.orc_build.argtypes = [C.POINTER(BuildSpec), C.POINTER(Problem), dp, ip, dp, dp]
        L.orc_pattern.restype = i64
        L.orc_pattern.argtypes = [C.POINTER(Problem), ip, ip]
        L.orc_coloring.argtypes = [C.POINTER(Problem), ip]
        L.orc_assemble.argtypes = [C.POINTER(Problem), dp, dp, C.c_double, dp, dp, C.c_int32]
        L.orc_vg_eval_deriv.argtypes = [C.POINTER(Problem), i64, dp, dp, dp, dp]
        L.orc_ilu0_factor.argtypes = [i64, C.c_int32, ip, ip, dp]
        L.orc_ilu0_apply.argtypes = [i64, C.c_int32, ip, ip, dp, dp, dp]
        L.orc_spmv.argtypes = [i64, C.c_int32, ip, ip, dp, dp, dp]
        L.orc_bicgstab.argtypes = [i64, C.c_int32, ip, ip, dp, dp, dp, C.c_double, C.c_int32,
                                   C.c_int32, C.POINTER(SolveInfo)]
        L.orc_default_newton_params.argtypes = [C.POINTER(NewtonParams)]
        L.orc_newton.argtypes = [C.POINTER(Problem), dp, dp, C.c_double, C.POINTER(NewtonParams),
                                 C.POINTER(NewtonInfo)]
        L.orc_time_loop.argtypes = [C.POINTER(Problem), dp, C.c_int32, C.c_double,
                                    C.POINTER(NewtonParams), C.POINTER(NewtonInfo)]
        u8 = C.POINTER(C.c_uint8)
        L.orc_partial_colors.restype = i64
        L.orc_partial_colors.argtypes = [i64, C.c_int32, dp, dp, C.c_double, dp, u8]
        L.orc_assemble_partial.argtypes = [C.POINTER(Problem), dp, dp, C.c_double, u8, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


def dpow(x, y):
    return lib().orc_dpow(float(x), float(y))


def dlog(x):
    return lib().orc_dlog(float(x))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def lognormal_K(seed, pid, n, mu, sigma):
    K = np.empty(n, dtype=np.float64)
    lib().orc_lognormal_K(seed, pid, n, mu, sigma, _dp(K))
    return K


class OracleProblem:
    """A problem description plus the arrays it points to (kept alive here)."""

    def __init__(self, kind, dim, n, seed=1, pid=0, mu_log10K=-11.0, sigma_log10K=0.5,
                 gravity=1, p_left=2e5, p_right=1e5, p_ref=1e5, influx=-0.04,
                 inlet_lo=0.25, inlet_hi=0.75, diff_method=0):
        nn = list(n) + [1] * (3 - len(n))
        self.spec = BuildSpec(kind=kind, dim=dim, n=(C.c_int32 * 3)(*nn), seed=seed, pid=pid,
                              mu_log10K=mu_log10K, sigma_log10K=sigma_log10K, gravity=gravity,
                              p_left=p_left, p_right=p_right, p_ref=p_ref, influx=influx,
                              inlet_lo=inlet_lo, inlet_hi=inlet_hi)
        N = int(np.prod(nn[:dim]))
        m = 2 if kind == 2 else 1
        nbf = lib().orc_num_boundary_faces(dim, (C.c_int32 * 3)(*nn))
        self.N, self.m, self.nbf = N, m, nbf
        self.K = np.empty(N)
        self.bc_type = np.empty(nbf, dtype=np.int32)
        self.bc_value = np.empty(nbf * m)
        self.u0 = np.empty(N * m)
        self.P = Problem()
        lib().orc_build(C.byref(self.spec), C.byref(self.P), _dp(self.K), _ip(self.bc_type),
                        _dp(self.bc_value), _dp(self.u0))
        self.P.diff_method = diff_method

    def set_materials(self, material_id, **material1):
        """Element-wise materials: material_id[N] in {0, 1}; material1 holds the van Genuchten
        parameters of material 1 (vg_alpha, vg_n, vg_l, swr, snr, thresholds), defaults from
        material 0."""
        self.material_id = np.ascontiguousarray(material_id, dtype=np.int32)
        self.P.material_id = _ip(self.material_id)
        for k in ("vg_alpha", "vg_n", "vg_l", "swr", "snr", "pc_low_swe", "pc_high_swe", "krn_low_swe",
                  "krw_high_swe"):
            setattr(self.P.material1, k, material1.get(k, getattr(self.P, k)))

    def set_arrays(self, K=None, bc_type=None, bc_value=None):
        if K is not None:
            self.K = np.ascontiguousarray(K, dtype=np.float64)
        if bc_type is not None:
            self.bc_type = np.ascontiguousarray(bc_type, dtype=np.int32)
        if bc_value is not None:
            self.bc_value = np.ascontiguousarray(bc_value, dtype=np.float64)
        self.P.K = _dp(self.K)
        self.P.bc_type = _ip(self.bc_type)
        self.P.bc_value = _dp(self.bc_value)

    def material_params(self):
        out = np.empty(24)
        lib().orc_material_params(C.byref(self.P), _dp(out))
        return out

    def pattern(self):
        nnzb = lib().orc_pattern(C.byref(self.P), None, None)
        rowptr = np.empty(self.N + 1, dtype=np.int32)
        colidx = np.empty(nnzb, dtype=np.int32)
        lib().orc_pattern(C.byref(self.P), _ip(rowptr), _ip(colidx))
        return rowptr, colidx

    def coloring(self):
        colors = np.empty(self.N, dtype=np.int32)
        nc = lib().orc_coloring(C.byref(self.P), _ip(colors))
        return nc, colors

    def assemble(self, u, uprev=None, dt=1.0, jacobian=True, nthreads=1, diff_method=None):
        """r (N*m) and J (nnzb*m*m); diff_method (default: the problem's) 0 numeric, 1 analytic."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.empty(self.N * self.m)
        J = None
        if jacobian:
            rowptr, colidx = self.pattern()
            J = np.empty(len(colidx) * self.m * self.m)
        up = None if uprev is None else np.ascontiguousarray(uprev, dtype=np.float64)
        keep = self.P.diff_method
        if diff_method is not None:
            self.P.diff_method = diff_method
        try:
            lib().orc_assemble(C.byref(self.P), _dp(u), _dp(up), dt, _dp(r), _dp(J), nthreads)
        finally:
            self.P.diff_method = keep
        return r, J

    def vg_deriv(self, sw):
        """Regularized dpc/dsw, dkrw/dsw, dkrn/dsw at absolute sw."""
        sw = np.ascontiguousarray(sw, dtype=np.float64)
        out = [np.zeros(sw.size) for _ in range(3)]
        lib().orc_vg_eval_deriv(C.byref(self.P), sw.size, _dp(sw), *[_dp(o) for o in out])
        return out

    def newton(self, u, uprev=None, dt=1.0, **kw):
        u = np.ascontiguousarray(u, dtype=np.float64).copy()
        p = default_newton_params(**kw)
        info = NewtonInfo()
        up = None if uprev is None else np.ascontiguousarray(uprev, dtype=np.float64)
        if lib().orc_newton(C.byref(self.P), _dp(u), _dp(up), dt, C.byref(p), C.byref(info)) == 2:
            raise ValueError("partial reassembly needs the numeric Jacobian")
        return u, info

    def time_loop(self, u, nsteps, dt, **kw):
        u = np.ascontiguousarray(u, dtype=np.float64).copy()
        p = default_newton_params(**kw)
        info = NewtonInfo()
        if lib().orc_time_loop(C.byref(self.P), _dp(u), nsteps, dt, C.byref(p), C.byref(info)) == 2:
            raise ValueError("partial reassembly needs the numeric Jacobian")
        return u, info

    def assemble_partial(self, u, uprev, dt, green, J):
        """assembleJacobianAndResidual with a partial reassembler's colors (green: N uint8, 1 keeps
        the column): returns r and J updated from the given previous Jacobian."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        up = None if uprev is None else np.ascontiguousarray(uprev, dtype=np.float64)
        green = np.ascontiguousarray(green, dtype=np.uint8)
        J = np.array(J, dtype=np.float64, copy=True)
        r = np.empty(self.N * self.m)
        if lib().orc_assemble_partial(C.byref(self.P), _dp(u), _dp(up), dt,
                                      green.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(r), _dp(J)) != 0:
            raise ValueError("partial reassembly needs the numeric Jacobian")
        return r, J


def partial_colors(u_last, u, m, threshold, dist):
    """The partial reassembler's colors after a Newton update (green: 1) and the updated
    distances; returns (green, dist, number of green cells)."""
    u_last = np.ascontiguousarray(u_last, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    dist = np.array(dist, dtype=np.float64, copy=True)
    green = np.zeros(dist.size, dtype=np.uint8)
    ng = lib().orc_partial_colors(dist.size, m, _dp(u_last), _dp(u), float(threshold), _dp(dist),
                                  green.ctypes.data_as(C.POINTER(C.c_uint8)))
    return green, dist, int(ng)


def default_newton_params(**kw):
    p = NewtonParams()
    lib().orc_default_newton_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def bicgstab(rowptr, colidx, A, b, m, reduction=1e-10, maxit=250, precond=1):
    x = np.zeros_like(b)
    info = SolveInfo()
    lib().orc_bicgstab(len(rowptr) - 1, m, _ip(rowptr), _ip(colidx), _dp(A), _dp(x), _dp(b),
                       reduction, maxit, precond, C.byref(info))
    return x, info


def ilu0(rowptr, colidx, A, m):
    lu = np.array(A, dtype=np.float64, copy=True)
    rc = lib().orc_ilu0_factor(len(rowptr) - 1, m, _ip(rowptr), _ip(colidx), _dp(lu))
    if rc != 0:
        raise RuntimeError("ILU0 failed")
    return lu


def ilu0_apply(rowptr, colidx, lu, d, m):
    v = np.zeros_like(d)
    lib().orc_ilu0_apply(len(rowptr) - 1, m, _ip(rowptr), _ip(colidx), _dp(lu), _dp(v), _dp(d))
    return v


def spmv(rowptr, colidx, A, x, m):
    y = np.empty_like(x)
    lib().orc_spmv(len(rowptr) - 1, m, _ip(rowptr), _ip(colidx), _dp(A), _dp(x), _dp(y))
    return y
