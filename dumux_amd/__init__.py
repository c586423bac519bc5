"""dumux_amd: MI355X (gfx950) CCTpfa Newton path of DuMux behind a C ABI.

This module is a thin ctypes binding of dumux_amd/build/libdumux_amd.so (see
include/dumux_amd.h). The classes carry the names of the reference interfaces they stand in
for (FVAssembler, ILUBiCGSTABSolver, NewtonSolver), with the reference's argument meaning:
the Jacobian and residual live on the GPU, the linear solver solves J du = r, Newton updates
u <- u - du. Errors: DMX_NUMERICAL_PROBLEM raises NumericalProblem (recoverable, as
Dumux::NumericalProblem), other codes raise RuntimeError.

The library is required: importing works without it, but every constructor raises if the
HIP library is missing; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libdumux_amd.so")

DMX_OK, DMX_NUMERICAL_PROBLEM = 0, 1
MODEL_ONEP, MODEL_TWOP = 1, 2
PRECOND_ILU0, PRECOND_JACOBI, PRECOND_AMG = 1, 2, 3
PRECISION_FP64, PRECISION_MIXED = 0, 1
DIFF_NUMERIC, DIFF_ANALYTIC = 0, 1
(KERNEL_ASSEMBLE, KERNEL_CELL_STATE, KERNEL_RESIDUAL, KERNEL_SPMV_FP64, KERNEL_SPMV_FP32, KERNEL_ILU0_APPLY_FP64,
 KERNEL_JACOBI_APPLY_FP32, KERNEL_BICG_HALF_FP32, KERNEL_NEWTON_UPDATE, KERNEL_REFINE_RESIDUAL,
 KERNEL_BJ_SPMV_DOT_FP32, KERNEL_BJ_SPMV_DOTS_FP32, KERNEL_BJ_ITERATION_FP32, KERNEL_MG_DOWN0, KERNEL_MG_UP0) = range(1, 16)


class NumericalProblem(RuntimeError):
    """Recoverable failure (dumux/common/exceptions.hh:26)."""


class VGMaterial(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("vg_alpha", "vg_n", "vg_l", "swr", "snr", "pc_low_swe", "pc_high_swe",
                                          "krn_low_swe", "krw_high_swe")]


class ProblemDesc(C.Structure):
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


class LinearParams(C.Structure):
    _fields_ = [
        ("reduction", C.c_double), ("max_iter", C.c_int32), ("precond", C.c_int32),
        ("precision", C.c_int32), ("inner_reduction", C.c_double), ("max_refinements", C.c_int32),
        ("ilu_sweep", C.c_int32),
    ]


class SolveInfo(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iterations", C.c_int32), ("refinements", C.c_int32),
                ("reduction", C.c_double), ("elapsed", C.c_double)]


class NewtonParams(C.Structure):
    _fields_ = [
        ("max_rel_shift", C.c_double), ("min_steps", C.c_int32), ("max_steps", C.c_int32),
        ("max_dt_divisions", C.c_int32), ("retry_factor", C.c_double),
        ("linear_pde", C.c_int32), ("verbosity", C.c_int32), ("linear", LinearParams),
        ("adaptive_dt", C.c_int32), ("target_steps", C.c_int32), ("max_dt", C.c_double),
        ("partial_reassembly", C.c_int32), ("reassembly_min_threshold", C.c_double),
        ("reassembly_max_threshold", C.c_double), ("reassembly_shift_weight", C.c_double),
    ]


class NewtonInfo(C.Structure):
    _fields_ = [
        ("converged", C.c_int32), ("newton_steps", C.c_int32), ("wasted_steps", C.c_int32),
        ("linear_iterations", C.c_int32), ("refinements", C.c_int32), ("time_steps", C.c_int32),
        ("dt_retries", C.c_int32), ("last_shift", C.c_double), ("t_assemble", C.c_double),
        ("t_solve", C.c_double), ("t_update", C.c_double), ("fp64_fallback", C.c_int32),
        ("reassembled_cells", C.c_int64),
    ]


class BatchItem(C.Structure):
    _fields_ = [("nsteps", C.c_int32), ("linear_pde", C.c_int32), ("dt", C.c_double), ("reduction", C.c_double)]


_lib = None

# every symbol declared in include/dumux_amd.h
EXPORTS = [
    "dmx_num_boundary_faces", "dmx_build_problem", "dmx_lognormal_permeability", "dmx_create",
    "dmx_destroy", "dmx_last_error", "dmx_num_dofs", "dmx_pattern", "dmx_level_schedule",
    "dmx_set_solution", "dmx_get_solution", "dmx_set_prev_solution", "dmx_reset_time_step",
    "dmx_advance_time_step", "dmx_set_time_step_size", "dmx_assemble_jacobian_and_residual",
    "dmx_assemble_residual", "dmx_get_residual", "dmx_get_jacobian", "dmx_default_linear_params",
    "dmx_linear_solve", "dmx_get_update", "dmx_residual_norm", "dmx_apply_preconditioner", "dmx_newton_update",
    "dmx_default_newton_params", "dmx_newton_solve", "dmx_time_loop", "dmx_material_params",
    "dmx_synchronize", "dmx_time_kernel", "dmx_batch_create", "dmx_batch_destroy",
    "dmx_batch_last_error", "dmx_batch_add", "dmx_batch_remove", "dmx_batch_size", "dmx_batch_step",
    "dmx_batch_status", "dmx_batch_run", "dmx_device_bytes", "dmx_estimate_bytes", "dmx_batch_time_kernel",
    "dmx_batch_autotune", "dmx_batch_set_tuning", "dmx_batch_get_tuning", "dmx_get_volume_variables",
    "dmx_partial_reassembly_reset", "dmx_partial_reassembly_update", "dmx_get_partial_colors",
    "dmx_assemble_jacobian_and_residual_partial",
]


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"dumux_amd HIP library not built: {LIB_PATH} (python dumux_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        dp, ip, i64, vp = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_int64, C.c_void_p
        sig = {
            "dmx_num_boundary_faces": (i64, [C.c_int32, ip]),
            "dmx_build_problem": (C.c_int, [C.POINTER(BuildSpec), C.POINTER(ProblemDesc), dp, ip, dp, dp]),
            "dmx_lognormal_permeability": (C.c_int, [C.c_uint32, C.c_uint32, i64, C.c_double, C.c_double, dp]),
            "dmx_create": (C.c_int, [C.POINTER(ProblemDesc), C.c_int32, C.POINTER(vp)]),
            "dmx_destroy": (C.c_int, [vp]),
            "dmx_last_error": (C.c_char_p, [vp]),
            "dmx_num_dofs": (C.c_int, [vp, C.POINTER(i64), ip, C.POINTER(i64)]),
            "dmx_pattern": (C.c_int, [vp, ip, ip]),
            "dmx_level_schedule": (C.c_int, [vp, ip, ip, ip]),
            "dmx_set_solution": (C.c_int, [vp, dp]),
            "dmx_get_solution": (C.c_int, [vp, dp]),
            "dmx_get_volume_variables": (C.c_int, [vp, dp, dp]),
            "dmx_set_prev_solution": (C.c_int, [vp, dp]),
            "dmx_reset_time_step": (C.c_int, [vp]),
            "dmx_advance_time_step": (C.c_int, [vp]),
            "dmx_set_time_step_size": (C.c_int, [vp, C.c_double]),
            "dmx_assemble_jacobian_and_residual": (C.c_int, [vp]),
            "dmx_assemble_residual": (C.c_int, [vp]),
            "dmx_get_residual": (C.c_int, [vp, dp]),
            "dmx_get_jacobian": (C.c_int, [vp, dp]),
            "dmx_default_linear_params": (None, [C.POINTER(LinearParams)]),
            "dmx_linear_solve": (C.c_int, [vp, C.POINTER(LinearParams), C.POINTER(SolveInfo)]),
            "dmx_get_update": (C.c_int, [vp, dp]),
            "dmx_residual_norm": (C.c_int, [vp, dp]),
            "dmx_apply_preconditioner": (C.c_int, [vp, C.c_int32, C.c_int32, dp, dp]),
            "dmx_newton_update": (C.c_int, [vp, dp]),
            "dmx_default_newton_params": (None, [C.POINTER(NewtonParams)]),
            "dmx_newton_solve": (C.c_int, [vp, C.POINTER(NewtonParams), C.POINTER(NewtonInfo)]),
            "dmx_time_loop": (C.c_int, [vp, C.c_int32, C.c_double, C.POINTER(NewtonParams), C.POINTER(NewtonInfo)]),
            "dmx_material_params": (C.c_int, [C.POINTER(ProblemDesc), dp]),
            "dmx_synchronize": (C.c_int, [vp]),
            "dmx_time_kernel": (C.c_int, [vp, C.c_int32, C.c_int32, dp, dp]),
            "dmx_batch_create": (C.c_int, [C.c_int32, C.POINTER(NewtonParams), C.POINTER(vp)]),
            "dmx_batch_destroy": (C.c_int, [vp]),
            "dmx_batch_last_error": (C.c_char_p, [vp]),
            "dmx_batch_add": (C.c_int, [vp, vp, C.POINTER(BatchItem)]),
            "dmx_batch_remove": (C.c_int, [vp, vp]),
            "dmx_batch_size": (C.c_int, [vp]),
            "dmx_batch_step": (C.c_int, [vp, C.c_int32, ip]),
            "dmx_batch_status": (C.c_int, [vp, vp, ip, C.POINTER(NewtonInfo)]),
            "dmx_batch_run": (C.c_int, [vp]),
            "dmx_device_bytes": (C.c_int, [vp, C.POINTER(i64)]),
            "dmx_batch_time_kernel": (C.c_int, [vp, C.c_int32, C.c_int32, dp, dp]),
            "dmx_batch_autotune": (C.c_int, [vp, C.c_int32, dp]),
            "dmx_batch_set_tuning": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32]),
            "dmx_batch_get_tuning": (C.c_int, [vp, ip, ip, ip]),
            "dmx_estimate_bytes": (C.c_int, [C.POINTER(ProblemDesc), C.c_int32, C.POINTER(i64)]),
            "dmx_partial_reassembly_reset": (C.c_int, [vp]),
            "dmx_partial_reassembly_update": (C.c_int, [vp, C.c_double, C.POINTER(i64)]),
            "dmx_get_partial_colors": (C.c_int, [vp, C.POINTER(C.c_uint8)]),
            "dmx_assemble_jacobian_and_residual_partial": (C.c_int, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


class Problem:
    """A problem description plus the host arrays it points to."""

    def __init__(self, kind, dim, n, seed=1, pid=0, mu_log10K=-11.0, sigma_log10K=0.5, gravity=1,
                 p_left=2e5, p_right=1e5, p_ref=1e5, influx=-0.04, inlet_lo=0.25, inlet_hi=0.75,
                 diff_method=DIFF_NUMERIC):
        """diff_method: DIFF_NUMERIC (forward differences) or DIFF_ANALYTIC (hand-coded
        derivatives), the DiffMethod template argument of the reference's FVAssembler."""
        nn = list(n) + [1] * (3 - len(n))
        self.spec = BuildSpec(kind=kind, dim=dim, n=(C.c_int32 * 3)(*nn), seed=seed, pid=pid,
                              mu_log10K=mu_log10K, sigma_log10K=sigma_log10K, gravity=gravity,
                              p_left=p_left, p_right=p_right, p_ref=p_ref, influx=influx,
                              inlet_lo=inlet_lo, inlet_hi=inlet_hi)
        L = lib()
        self.N = int(np.prod(nn[:dim]))
        self.m = 2 if kind == 2 else 1
        self.nbf = L.dmx_num_boundary_faces(dim, (C.c_int32 * 3)(*nn))
        self.K = np.empty(self.N)
        self.bc_type = np.empty(self.nbf, dtype=np.int32)
        self.bc_value = np.empty(self.nbf * self.m)
        self.u0 = np.empty(self.N * self.m)
        self.desc = ProblemDesc()
        rc = L.dmx_build_problem(C.byref(self.spec), C.byref(self.desc), _dp(self.K), _ip(self.bc_type),
                                 _dp(self.bc_value), _dp(self.u0))
        if rc != DMX_OK:
            raise RuntimeError(f"dmx_build_problem failed ({rc})")
        self.desc.diff_method = diff_method

    def set_arrays(self, K=None, bc_type=None, bc_value=None):
        if K is not None:
            self.K = np.ascontiguousarray(K, dtype=np.float64)
        if bc_type is not None:
            self.bc_type = np.ascontiguousarray(bc_type, dtype=np.int32)
        if bc_value is not None:
            self.bc_value = np.ascontiguousarray(bc_value, dtype=np.float64)
        self.desc.K = _dp(self.K)
        self.desc.bc_type = _ip(self.bc_type)
        self.desc.bc_value = _dp(self.bc_value)

    def set_materials(self, material_id, **material1):
        """Element-wise materials (SpatialParams::fluidMatrixInteraction per element):
        material_id[N] in {0, 1}; material1 holds the van Genuchten parameters of material 1
        (vg_alpha, vg_n, vg_l, swr, snr and the regularization thresholds; defaults: material 0's)."""
        self.material_id = np.ascontiguousarray(material_id, dtype=np.int32)
        self.desc.material_id = _ip(self.material_id)
        for k in ("vg_alpha", "vg_n", "vg_l", "swr", "snr", "pc_low_swe", "pc_high_swe", "krn_low_swe",
                  "krw_high_swe"):
            setattr(self.desc.material1, k, material1.get(k, getattr(self.desc, k)))

    def estimate_bytes(self, batched=True, multigrid=False):
        """Device bytes of a context of this problem (batched: with the batch executor's arrays;
        multigrid: and the multigrid levels of a batch with PRECOND_AMG)."""
        v = C.c_int64()
        mode = (2 if multigrid else 1) if batched else 0
        if lib().dmx_estimate_bytes(C.byref(self.desc), mode, C.byref(v)) != DMX_OK:
            raise RuntimeError("dmx_estimate_bytes failed")
        return v.value

    def material_params(self):
        out = np.empty(24)
        lib().dmx_material_params(C.byref(self.desc), _dp(out))
        return out


def estimate_bytes(kind, dim, n, batched=True, multigrid=False):
    """Device bytes of a context of a kind-`kind` problem on an n-cell box, before creating it."""
    nn = list(n) + [1] * (3 - len(n))
    d = ProblemDesc()
    d.dim, d.model = dim, 2 if kind == 2 else 1
    d.n = (C.c_int32 * 3)(*nn)
    v = C.c_int64()
    mode = (2 if multigrid else 1) if batched else 0
    if lib().dmx_estimate_bytes(C.byref(d), mode, C.byref(v)) != DMX_OK:
        raise RuntimeError("dmx_estimate_bytes failed")
    return v.value


def _check(rc, ctx, what):
    if rc == DMX_OK:
        return
    msg = lib().dmx_last_error(ctx).decode() if ctx else ""
    if rc == DMX_NUMERICAL_PROBLEM:
        raise NumericalProblem(f"{what}: {msg}")
    raise RuntimeError(f"{what} failed ({rc}): {msg}")


class FVAssembler:
    """Device-resident stand-in for Dumux::FVAssembler (dumux/assembly/fvassembler.hh)."""

    def __init__(self, problem, device=0):
        self.problem = problem
        self.ctx = C.c_void_p()
        _check(lib().dmx_create(C.byref(problem.desc), device, C.byref(self.ctx)), None, "dmx_create")
        n, m, nnzb = C.c_int64(), C.c_int32(), C.c_int64()
        lib().dmx_num_dofs(self.ctx, C.byref(n), C.byref(m), C.byref(nnzb))
        self.N, self.m, self.nnzb = n.value, m.value, nnzb.value

    def close(self):
        if self.ctx:
            lib().dmx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def numDofs(self):
        return self.N

    def device_bytes(self):
        v = C.c_int64()
        _check(lib().dmx_device_bytes(self.ctx, C.byref(v)), self.ctx, "device_bytes")
        return v.value

    def pattern(self):
        rowptr = np.empty(self.N + 1, dtype=np.int32)
        colidx = np.empty(self.nnzb, dtype=np.int32)
        _check(lib().dmx_pattern(self.ctx, _ip(rowptr), _ip(colidx)), self.ctx, "pattern")
        return rowptr, colidx

    def level_schedule(self):
        nl = C.c_int32()
        lib().dmx_level_schedule(self.ctx, C.byref(nl), None, None)
        lp = np.empty(nl.value + 1, dtype=np.int32)
        lr = np.empty(self.N, dtype=np.int32)
        lib().dmx_level_schedule(self.ctx, C.byref(nl), _ip(lp), _ip(lr))
        return lp, lr

    def setSolution(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        _check(lib().dmx_set_solution(self.ctx, _dp(u)), self.ctx, "set_solution")

    def solution(self):
        u = np.empty(self.N * self.m)
        _check(lib().dmx_get_solution(self.ctx, _dp(u)), self.ctx, "get_solution")
        return u

    def volumeVariables(self):
        """(phase pressures, rho*mobility), each N*m (cell-major), of the current solution."""
        p, rm = np.empty(self.N * self.m), np.empty(self.N * self.m)
        _check(lib().dmx_get_volume_variables(self.ctx, _dp(p), _dp(rm)), self.ctx, "get_volume_variables")
        return p, rm

    def setPreviousSolution(self, uprev):
        up = None if uprev is None else np.ascontiguousarray(uprev, dtype=np.float64)
        _check(lib().dmx_set_prev_solution(self.ctx, _dp(up)), self.ctx, "set_prev_solution")

    def setTimeStepSize(self, dt):
        _check(lib().dmx_set_time_step_size(self.ctx, dt), self.ctx, "set_time_step_size")

    def resetTimeStep(self):
        _check(lib().dmx_reset_time_step(self.ctx), self.ctx, "reset_time_step")

    def advanceTimeStep(self):
        _check(lib().dmx_advance_time_step(self.ctx), self.ctx, "advance_time_step")

    def assembleJacobianAndResidual(self, u=None):
        if u is not None:
            self.setSolution(u)
        _check(lib().dmx_assemble_jacobian_and_residual(self.ctx), self.ctx, "assembleJacobianAndResidual")

    def assembleResidual(self, u=None):
        if u is not None:
            self.setSolution(u)
        _check(lib().dmx_assemble_residual(self.ctx), self.ctx, "assembleResidual")

    def newtonUpdate(self):
        """u <- uLastIter - deltaU (newtonsolver.hh:543-604); returns the maximum relative shift."""
        v = C.c_double()
        _check(lib().dmx_newton_update(self.ctx, C.byref(v)), self.ctx, "newtonUpdate")
        return v.value

    # the partial reassembler (dumux/assembly/partialreassembler.hh), kept on the device
    def resetPartialReassembly(self):
        """resetColors and zero distances (newtonBegin): every cell red."""
        _check(lib().dmx_partial_reassembly_reset(self.ctx), self.ctx, "resetColors")

    def updatePartialReassembly(self, threshold):
        """After newtonUpdate: distances and colors for the threshold; returns the green cells."""
        ng = C.c_int64()
        _check(lib().dmx_partial_reassembly_update(self.ctx, float(threshold), C.byref(ng)), self.ctx, "computeColors")
        return ng.value

    def partialColors(self):
        """dofColor of every cell: 1 green (the Jacobian column is kept), 0 red."""
        g = np.empty(self.N, dtype=np.uint8)
        _check(lib().dmx_get_partial_colors(self.ctx, g.ctypes.data_as(C.POINTER(C.c_uint8))), self.ctx, "colors")
        return g

    def assembleJacobianAndResidualPartial(self, u=None):
        """assembleJacobianAndResidual(u, partialReassembler): green cells keep their columns."""
        if u is not None:
            self.setSolution(u)
        _check(lib().dmx_assemble_jacobian_and_residual_partial(self.ctx), self.ctx, "assembleJacobianAndResidual")

    def residual(self):
        r = np.empty(self.N * self.m)
        _check(lib().dmx_get_residual(self.ctx, _dp(r)), self.ctx, "residual")
        return r

    def jacobian(self):
        J = np.empty(self.nnzb * self.m * self.m)
        _check(lib().dmx_get_jacobian(self.ctx, _dp(J)), self.ctx, "jacobian")
        return J

    def update(self):
        du = np.empty(self.N * self.m)
        _check(lib().dmx_get_update(self.ctx, _dp(du)), self.ctx, "update")
        return du

    def synchronize(self):
        _check(lib().dmx_synchronize(self.ctx), self.ctx, "synchronize")

    def time_kernel(self, kernel, reps=20):
        """(average ms, algorithmic bytes) of one kernel launch, HIP events on the context stream."""
        ms, b = C.c_double(), C.c_double()
        _check(lib().dmx_time_kernel(self.ctx, kernel, reps, C.byref(ms), C.byref(b)), self.ctx, "time_kernel")
        return ms.value, b.value


def linear_params(**kw):
    p = LinearParams()
    lib().dmx_default_linear_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class ILUBiCGSTABSolver:
    """Stand-in for Dumux::ILUBiCGSTABIstlSolver (dumux/linear/istlsolvers.hh:635-642)."""

    def __init__(self, **kw):
        kw.setdefault("precond", PRECOND_ILU0)
        self.params = linear_params(**kw)

    def setResidualReduction(self, r):
        self.params.reduction = r

    def solve(self, assembler):
        info = SolveInfo()
        rc = lib().dmx_linear_solve(assembler.ctx, C.byref(self.params), C.byref(info))
        if rc < 0:
            _check(rc, assembler.ctx, "linear solve")
        return info

    def norm(self, assembler):
        v = C.c_double()
        _check(lib().dmx_residual_norm(assembler.ctx, C.byref(v)), assembler.ctx, "norm")
        return v.value

    def applyPreconditioner(self, assembler, d):
        """v = M^-1 d once, with this solver's preconditioner built on the assembled Jacobian
        (in the solver's precision: mixed rounds d to FP32 first)."""
        d = np.ascontiguousarray(d, dtype=np.float64)
        v = np.empty_like(d)
        _check(lib().dmx_apply_preconditioner(assembler.ctx, self.params.precond, self.params.precision, _dp(d), _dp(v)),
               assembler.ctx, "preconditioner")
        return v


class AMGBiCGSTABSolver(ILUBiCGSTABSolver):
    """Stand-in for Dumux::AMGBiCGSTABIstlSolver (dumux/linear/istlsolvers.hh:729-736): BiCGSTAB
    preconditioned by the aggregation multigrid V-cycle."""

    def __init__(self, **kw):
        kw.setdefault("precond", PRECOND_AMG)
        super().__init__(**kw)


def newton_params(linear=None, **kw):
    p = NewtonParams()
    lib().dmx_default_newton_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    if linear:
        for k, v in linear.items():
            setattr(p.linear, k, v)
    return p


class NewtonSolver:
    """Stand-in for Dumux::NewtonSolver (dumux/nonlinear/newtonsolver.hh) driving the device
    assembler; the Newton loop itself runs in the library's C++ host code."""

    def __init__(self, assembler, linear=None, **kw):
        self.assembler = assembler
        self.params = newton_params(linear=linear, **kw)

    def solve(self, dt=None):
        if dt is not None:
            self.assembler.setTimeStepSize(dt)
        info = NewtonInfo()
        rc = lib().dmx_newton_solve(self.assembler.ctx, C.byref(self.params), C.byref(info))
        if rc < 0:
            _check(rc, self.assembler.ctx, "newton")
        return info

    def time_loop(self, nsteps, dt):
        info = NewtonInfo()
        rc = lib().dmx_time_loop(self.assembler.ctx, nsteps, dt, C.byref(self.params), C.byref(info))
        if rc < 0:
            _check(rc, self.assembler.ctx, "time loop")
        return info


class Batch:
    """Batched executor with iteration-level continuous batching: many FVAssembler contexts
    advanced together, each by its own device-side Newton / refinement state machine (FP32
    block-Jacobi BiCGSTAB in FP64 iterative refinement). Problems join with add() and leave with
    remove() between step() calls; each keeps its own time level, dt control and Newton state."""

    def __init__(self, device=0, linear=None, **newton_kw):
        lin = dict(precond=PRECOND_JACOBI, precision=PRECISION_MIXED, reduction=1e-10, inner_reduction=1e-4,
                   max_iter=2000, max_refinements=30)
        lin.update(linear or {})
        self.params = newton_params(linear=lin, **newton_kw)
        self.b = C.c_void_p()
        rc = lib().dmx_batch_create(device, C.byref(self.params), C.byref(self.b))
        if rc != DMX_OK:
            raise RuntimeError(f"dmx_batch_create failed ({rc})")
        self.members = []

    def _check(self, rc, what):
        if rc != DMX_OK:
            raise RuntimeError(f"{what} failed ({rc}): {lib().dmx_batch_last_error(self.b).decode()}")

    def add(self, assembler, nsteps=1, dt=1.0, linear_pde=False, reduction=0.0):
        item = BatchItem(nsteps, 1 if linear_pde else 0, dt, reduction)
        self._check(lib().dmx_batch_add(self.b, assembler.ctx, C.byref(item)), "batch add")
        self.members.append(assembler)

    def remove(self, assembler):
        self._check(lib().dmx_batch_remove(self.b, assembler.ctx), "batch remove")
        self.members.remove(assembler)

    def __len__(self):
        return lib().dmx_batch_size(self.b)

    def step(self, ticks=1):
        """Run `ticks` ticks; returns the number of members still running."""
        n = C.c_int32()
        self._check(lib().dmx_batch_step(self.b, ticks, C.byref(n)), "batch step")
        return n.value

    def run(self):
        self._check(lib().dmx_batch_run(self.b), "batch run")

    def time_kernel(self, kernel, reps=20):
        """(average ms, algorithmic bytes summed over the members) of one batched launch."""
        ms, b = C.c_double(), C.c_double()
        self._check(lib().dmx_batch_time_kernel(self.b, kernel, reps, C.byref(ms), C.byref(b)), "batch time_kernel")
        return ms.value, b.value

    def autotune(self, reps=10):
        """Pick the fastest launch shape of the BiCGSTAB kernels for the current members;
        returns (coarse, vector grid, SpMV grid, microseconds per iteration)."""
        us = C.c_double()
        self._check(lib().dmx_batch_autotune(self.b, reps, C.byref(us)), "batch autotune")
        return self.tuning() + (us.value,)

    def tuning(self):
        c, v, s = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(lib().dmx_batch_get_tuning(self.b, C.byref(c), C.byref(v), C.byref(s)), "batch get_tuning")
        return c.value, v.value, s.value

    def set_tuning(self, coarse, vec_grid, spmv_grid):
        self._check(lib().dmx_batch_set_tuning(self.b, coarse, vec_grid, spmv_grid), "batch set_tuning")

    def status(self, assembler):
        """(finished, NewtonInfo): finished 1 = reached its end time, -1 = failed, 0 = running."""
        f, info = C.c_int32(), NewtonInfo()
        self._check(lib().dmx_batch_status(self.b, assembler.ctx, C.byref(f), C.byref(info)), "batch status")
        return f.value, info

    def close(self):
        if self.b:
            lib().dmx_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
