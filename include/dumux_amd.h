/* dumux_amd: MI355X (gfx950) implementation of DuMux's cell-centred TPFA Newton path for
 * porous-medium flow (OneP and TwoP immiscible, pw-Sn), behind a C ABI.
 *
 * The entry points replace, one to one, the member functions that DuMux's NewtonSolver
 * calls on its assembler and linear solver:
 *   FVAssembler            dumux/assembly/fvassembler.hh
 *   IstlIterativeLinearSolver / ILUBiCGSTABIstlSolver   dumux/linear/istlsolvers.hh
 *   NewtonSolver           dumux/nonlinear/newtonsolver.hh
 * Plain pointers and sizes only. Host pointers unless a name says _dev. Every function returns
 * an int status: DMX_OK, DMX_NUMERICAL_PROBLEM (recoverable, the reference throws
 * Dumux::NumericalProblem, dumux/common/exceptions.hh:26) or a negative fatal code. No C++
 * exception crosses the ABI. A context is single-threaded: one HIP stream per context;
 * contexts may share a device.
 *
 * Layouts (the reference's): solution / residual vectors are BlockVector<FieldVector<double,m>>
 * (N*m doubles, the m unknowns of a cell contiguous); the Jacobian is
 * BCRSMatrix<FieldMatrix<double,m,m>> (dumux/discretization/fvproperties.hh:62-75): block row
 * pointer (N+1 int32), sorted block column indices (nnzb int32), blocks m*m row-major.
 */
#ifndef DUMUX_AMD_H
#define DUMUX_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DMX_OK = 0,
    DMX_NUMERICAL_PROBLEM = 1,   /* linear solver not converged / non-finite values */
    DMX_ERR_INVALID = -1,        /* bad argument or state (Dune::InvalidStateException) */
    DMX_ERR_DEVICE = -2,         /* HIP runtime error */
    DMX_ERR_NOMEM = -3,          /* device allocation failed */
    DMX_ERR_UNSUPPORTED = -4
};

enum { DMX_MODEL_ONEP = 1, DMX_MODEL_TWOP = 2 };
/* DMX_PRECOND_AMG: aggregation multigrid (the role of AMGBiCGSTABIstlSolver,
 * linear/istlsolvers.hh:729-736) on 2x2x2 cell aggregates with Galerkin coarse operators of the
 * unscaled Jacobian, one damped block-Jacobi step before and after each coarse correction, as a
 * right preconditioner of BiCGSTAB: in the batched executor (FP32 inside FP64 refinement) and in
 * dmx_linear_solve (FP64 or mixed precision; DMX_ERR_UNSUPPORTED for grids without a coarse level) */
enum { DMX_PRECOND_ILU0 = 1, DMX_PRECOND_JACOBI = 2, DMX_PRECOND_AMG = 3 };
enum { DMX_PRECISION_FP64 = 0, DMX_PRECISION_MIXED = 1 };
/* FVAssembler<TypeTag, DiffMethod> (dumux/assembly/fvassembler.hh:58-63): numeric = forward
 * differences (CCLocalAssembler<numeric>, cclocalassembler.hh:164-348); analytic = hand-coded
 * derivatives of the incompressible 1p / 2p local residuals (CCLocalAssembler<analytic>,
 * cclocalassembler.hh:490-580). */
enum { DMX_DIFF_NUMERIC = 0, DMX_DIFF_ANALYTIC = 1 };

/* van Genuchten law of one material (VanGenuchtenDefault, vangenuchten.hh:629-630) */
typedef struct dmx_vg_material {
    double vg_alpha, vg_n, vg_l, swr, snr;
    double pc_low_swe, pc_high_swe, krn_low_swe, krw_high_swe;
} dmx_vg_material;

/* Problem on a structured dim-D box, cells x-fastest, isotropic permeability per cell.
 * Boundary faces are numbered side by side (-x,+x,-y,+y,-z,+z) and inside a side by the two
 * tangential cell coordinates, lower axis fastest. This carries what the reference obtains
 * from Problem::boundaryTypes / dirichlet / neumann (dumux/common/fvproblem.hh:126-516),
 * SpatialParams::permeability / porosity / fluidMatrixInteraction and the FluidSystem. */
typedef struct dmx_problem_desc {
    int32_t dim;            /* 2 or 3 */
    int32_t model;          /* DMX_MODEL_ONEP or DMX_MODEL_TWOP (p0s1: pw, Sn; water wetting) */
    int32_t n[3];           /* cells per axis (n[2] = 1 for dim 2) */
    int32_t gravity;        /* Problem.EnableGravity, g = -9.81 e_dim */
    double lower[3], upper[3];
    double rho[2], mu[2];   /* phase 0 = wetting (water), phase 1 = non-wetting */
    double porosity;
    double vg_alpha, vg_n, vg_l, swr, snr;                    /* VanGenuchtenDefault */
    double pc_low_swe, pc_high_swe, krn_low_swe, krw_high_swe; /* regularization thresholds */
    const double* K;        /* N permeabilities [m^2] */
    const int32_t* bc_type; /* per boundary face: 0 Neumann, 1 Dirichlet */
    const double* bc_value; /* per boundary face * m: Dirichlet primary variables or Neumann
                               flux per area [kg/(m^2 s)], negative = inflow */
    /* NumericEpsilon (dumux/assembly/numericepsilon.hh:37-53): eps = fd_base_eps*fd_magnitude[pv]
     * if fd_magnitude[pv] > 0, else fd_base_eps*(|x|+1). Reference defaults 1e-10 and -1. */
    double fd_base_eps;
    double fd_magnitude[2];
    int32_t diff_method;    /* DMX_DIFF_NUMERIC or DMX_DIFF_ANALYTIC */
    /* Element-wise materials (SpatialParams::fluidMatrixInteraction(element, ...), e.g. the lens
     * of test/porousmediumflow/2p/incompressible/spatialparams.hh:64-118): material_id[N] in
     * {0, 1} selects the vg_* law above (0) or material1 (1); NULL: every cell uses the vg_* law.
     * A Dirichlet ghost state uses its boundary cell's law (tpfa/elementvolumevariables.hh:326-338). */
    const int32_t* material_id;
    dmx_vg_material material1;
} dmx_problem_desc;

/* Synthetic problems of the benchmark workloads (seeded log-normal K):
 * kind 1: OneP, Dirichlet p_left on -x and p_right on +x, no-flow elsewhere, initial p_ref.
 * kind 2: TwoP DNAPL infiltration, hydrostatic Dirichlet pw (Sn = 0) on the lateral sides,
 *         DNAPL Neumann flux `influx` on the top face centre (inlet_lo, inlet_hi) fraction of
 *         each tangential axis, no-flow bottom, hydrostatic initial pw and Sn = 0.
 * Mirrors the callbacks of test/porousmediumflow/2p/incompressible/problem.hh:55-122. */
typedef struct dmx_build_spec {
    int32_t kind, dim;
    int32_t n[3];
    uint32_t seed, pid;
    double mu_log10K, sigma_log10K;
    int32_t gravity;
    double p_left, p_right;
    double p_ref;
    double influx;
    double inlet_lo, inlet_hi;
} dmx_build_spec;

typedef struct dmx_context dmx_context;

/* ---- problem setup ---------------------------------------------------------------------- */
int64_t dmx_num_boundary_faces(int32_t dim, const int32_t* n);
/* Fill desc and the caller-sized arrays K[N], bc_type[nbf], bc_value[nbf*m], u0[N*m]. */
int dmx_build_problem(const dmx_build_spec* spec, dmx_problem_desc* desc, double* K,
                      int32_t* bc_type, double* bc_value, double* u0);
/* Seeded log-normal permeability field (Philox4x32-10 + polar method), host side. */
int dmx_lognormal_permeability(uint32_t seed, uint32_t pid, int64_t n, double mu_log10K,
                               double sigma_log10K, double* K);

/* ---- FVAssembler (dumux/assembly/fvassembler.hh) ---------------------------------------
 * dmx_create ~ FVAssembler ctor + setLinearSystem() (fvassembler.hh:131-172, 294-302): builds
 * the BCRS pattern (jacobianpattern.hh:25-53), uploads K and the boundary data. */
int dmx_create(const dmx_problem_desc* desc, int32_t device, dmx_context** ctx);
int dmx_destroy(dmx_context* ctx);
const char* dmx_last_error(const dmx_context* ctx);
/* device memory held by a context, and the estimate of it before creation (batched 1: plus the
 * arrays the batched executor adds; 2: and its multigrid levels, DMX_PRECOND_AMG). No reference
 * counterpart: the admission gate of serving. */
int dmx_device_bytes(const dmx_context* ctx, int64_t* bytes);
int dmx_estimate_bytes(const dmx_problem_desc* desc, int32_t batched, int64_t* bytes);
/* numDofs() (fvassembler.hh:315): cells, unknowns per cell, BCRS blocks */
int dmx_num_dofs(const dmx_context* ctx, int64_t* ncells, int32_t* m, int64_t* nnzb);
/* the Jacobian pattern: rowptr[N+1], colidx[nnzb] */
int dmx_pattern(const dmx_context* ctx, int32_t* rowptr, int32_t* colidx);
/* ILU(0) level schedule: level_ptr[nlevels+1], level_rows[N] (rows of one level ascending) */
int dmx_level_schedule(const dmx_context* ctx, int32_t* nlevels, int32_t* level_ptr, int32_t* level_rows);

/* SolutionVector state. set_solution ~ the curSol argument of assembleJacobianAndResidual /
 * updateGridVariables (fvassembler.hh:179, 379); set_prev_solution ~ setPreviousSolution
 * (fvassembler.hh:361), NULL makes the problem stationary (isStationaryProblem, :367);
 * reset_time_step ~ resetTimeStep (:387) (u <- uprev); advance_time_step: uprev <- u. */
int dmx_set_solution(dmx_context* ctx, const double* u);
int dmx_get_solution(const dmx_context* ctx, double* u);
/* Volume variables of the current solution (GridVolumeVariables, as VtkOutputModule reads them,
 * dumux/io/vtkoutputmodule.hh): phase pressures p[N*m] (pw, pn = pw + pc) and rho*mobility
 * rho_mob[N*m] per cell (2p/volumevariables.hh:74-191, 1p/volumevariables.hh:64-120). */
int dmx_get_volume_variables(dmx_context* ctx, double* p, double* rho_mob);
int dmx_set_prev_solution(dmx_context* ctx, const double* uprev);
int dmx_reset_time_step(dmx_context* ctx);
int dmx_advance_time_step(dmx_context* ctx);
/* TimeLoop::timeStepSize (dumux/common/timeloop.hh:353) */
int dmx_set_time_step_size(dmx_context* ctx, double dt);

/* assembleJacobianAndResidual(curSol) (fvassembler.hh:178-207) on the device-resident
 * solution; numeric forward differences (cclocalassembler.hh:164-348). */
int dmx_assemble_jacobian_and_residual(dmx_context* ctx);
/* assembleResidual(curSol) (fvassembler.hh:240-267) */
int dmx_assemble_residual(dmx_context* ctx);
/* residual() / jacobian() (fvassembler.hh:339-344), copied to the host */
int dmx_get_residual(const dmx_context* ctx, double* r);
int dmx_get_jacobian(const dmx_context* ctx, double* values);

/* ---- linear solver (dumux/linear/istlsolvers.hh) --------------------------------------- */
typedef struct dmx_linear_params {
    double reduction;     /* LinearSolver.ResidualReduction (Newton sets it, newtonsolver.hh:232) */
    int32_t max_iter;     /* LinearSolver.MaxIterations (250) */
    int32_t precond;      /* DMX_PRECOND_ILU0 (ILUBiCGSTABIstlSolver) or DMX_PRECOND_JACOBI */
    int32_t precision;    /* DMX_PRECISION_FP64, or DMX_PRECISION_MIXED: FP32 Krylov inside FP64
                             iterative refinement */
    double inner_reduction; /* mixed: relative reduction of each FP32 solve */
    int32_t max_refinements;/* mixed: maximum FP64 refinement steps */
    int32_t ilu_sweep;      /* ILU(0) triangular sweeps (same bits either way; no counterpart in the
                               reference): 0 = one launch per sweep, rows drawn in level order;
                               1 = one launch per level */
} dmx_linear_params;

typedef struct dmx_solve_info {  /* IstlSolverResult (istlsolvers.hh:153-163) */
    int32_t converged;
    int32_t iterations;          /* Krylov iterations (summed over refinement steps) */
    int32_t refinements;         /* FP64 refinement steps (mixed precision) */
    double reduction;            /* final ||b - A x|| / ||b|| */
    double elapsed;              /* seconds */
} dmx_solve_info;

void dmx_default_linear_params(dmx_linear_params* p);
/* solve(A, x, b) (istlsolvers.hh:273): A = the assembled Jacobian, b = the residual,
 * x = the Newton update deltaU (zeroed first, newtonsolver.hh:1032). */
int dmx_linear_solve(dmx_context* ctx, const dmx_linear_params* p, dmx_solve_info* info);
int dmx_get_update(const dmx_context* ctx, double* du);
/* norm(b) (istlsolvers.hh:306-337) of the current residual */
int dmx_residual_norm(dmx_context* ctx, double* norm);
/* v = M^-1 d once, for the preconditioner the linear solver builds on the assembled Jacobian
 * (the preconditioner's apply inside IstlIterativeLinearSolver, istlsolvers.hh:535-568):
 * precond as in dmx_linear_params, precision DMX_PRECISION_FP64 (FP64 vectors; the multigrid
 * V-cycle itself runs in FP32) or DMX_PRECISION_MIXED (the FP32 operand of the inner solve,
 * rounded from d). d and v are host arrays of ncells * m values. ILU(0) sweeps run in the form
 * of the last dmx_linear_solve (dmx_linear_params.ilu_sweep; one launch per sweep before any). */
int dmx_apply_preconditioner(dmx_context* ctx, int32_t precond, int32_t precision, const double* d, double* v);

/* ---- PartialReassembler (dumux/assembly/partialreassembler.hh, CCTpfa engine :301-375) ----
 * Numeric (forward-difference) Jacobian only: the assembly owns columns, and a green cell keeps
 * its column of the previous Jacobian. The analytic assembler returns DMX_ERR_UNSUPPORTED (the
 * reference's analytic tests run with Newton.EnablePartialReassembly false). */
/* resetColors and the distances of NewtonSolver::newtonBegin (newtonsolver.hh:415-419): every
 * cell red, every distance 0 */
int dmx_partial_reassembly_reset(dmx_context* ctx);
/* after dmx_newton_update: each cell's distance from the last linearization grows by its maximum
 * relative shift; cells above threshold turn red (distance reset), the others green
 * (newtonsolver.hh:565-603, computeColors partialreassembler.hh:308-346); green_cells may be null */
int dmx_partial_reassembly_update(dmx_context* ctx, double threshold, int64_t* green_cells);
/* dofColor (partialreassembler.hh:366-370) of every cell: 1 green, 0 red */
int dmx_get_partial_colors(const dmx_context* ctx, uint8_t* green);
/* assembleJacobianAndResidual(curSol, partialReassembler) (fvassembler.hh:178-207, 434-445,
 * cclocalassembler.hh:61-76): red cells' Jacobian columns are recomputed, green cells keep
 * theirs and evaluate their residual only */
int dmx_assemble_jacobian_and_residual_partial(dmx_context* ctx);

/* ---- NewtonSolver (dumux/nonlinear/newtonsolver.hh) -------------------------------------
 * newtonUpdate (newtonsolver.hh:543-604): u <- uLastIter - deltaU, then the maximum relative
 * shift (:110-128, 1138-1144) between uLastIter and u. */
int dmx_newton_update(dmx_context* ctx, double* shift);

typedef struct dmx_newton_params {
    double max_rel_shift;     /* Newton.MaxRelativeShift 1e-8 */
    int32_t min_steps;        /* Newton.MinSteps 2 */
    int32_t max_steps;        /* Newton.MaxSteps 18 */
    int32_t max_dt_divisions; /* Newton.MaxTimeStepDivisions 10 */
    double retry_factor;      /* Newton.RetryTimeStepReductionFactor 0.5 */
    int32_t linear_pde;       /* 1: LinearPDESolver (assemble, solve, update once) */
    int32_t verbosity;
    dmx_linear_params linear;
    /* time loops: 1 = the next dt is suggestTimeStepSize of the last step (newtonsolver.hh:784-798,
     * as test/porousmediumflow/2p/incompressible/main.cc:162), capped at max_dt and the end time
     * (TimeLoop::setTimeStepSize, common/timeloop.hh:320-324, 404-411); 0 = fixed dt */
    int32_t adaptive_dt;
    int32_t target_steps;     /* Newton.TargetSteps 10 */
    double max_dt;            /* TimeLoop max time step size; <= 0: none */
    /* partial reassembly (newtonsolver.hh:565-603, 1241-1244): reassemble the Jacobian columns of
     * cells whose summed relative shift since their last linearization exceeds
     * max(min_threshold, min(max_threshold, shift_weight * shift)); single-problem solves with
     * the numeric Jacobian (the batched executor returns DMX_ERR_UNSUPPORTED) */
    int32_t partial_reassembly;      /* Newton.EnablePartialReassembly 0 */
    double reassembly_min_threshold; /* Newton.ReassemblyMinThreshold; <= 0: 1e-1 * max_rel_shift */
    double reassembly_max_threshold; /* Newton.ReassemblyMaxThreshold; <= 0: 1e2 * max_rel_shift */
    double reassembly_shift_weight;  /* Newton.ReassemblyShiftWeight 1e-3 */
} dmx_newton_params;

typedef struct dmx_newton_info {
    int32_t converged;
    int32_t newton_steps;     /* iterations of converged solves */
    int32_t wasted_steps;     /* iterations of failed attempts */
    int32_t linear_iterations;
    int32_t refinements;
    int32_t time_steps;
    int32_t dt_retries;
    double last_shift;
    double t_assemble, t_solve, t_update;  /* seconds (Dune::Timer split, newtonsolver.hh:950-956) */
    int32_t fp64_fallback;    /* batched executor: the FP32 refinement stalled, solves ran in FP64 */
    int64_t reassembled_cells;/* Jacobian columns computed, summed over the assemblies (the
                                 partial reassembler's report, partialreassembler.hh:458-468) */
} dmx_newton_info;

void dmx_default_newton_params(dmx_newton_params* p);
/* NewtonSolver::solve(vars) (newtonsolver.hh:363-369) at the current dt */
int dmx_newton_solve(dmx_context* ctx, const dmx_newton_params* p, dmx_newton_info* info);
/* implicit Euler from the current solution (also the first old time level, as xOld = x in the
 * reference's mains) to nsteps*dt with NewtonSolver::solve(vars, timeLoop) retries
 * (newtonsolver.hh:309-355); the solution stays on the device */
int dmx_time_loop(dmx_context* ctx, int32_t nsteps, double dt, const dmx_newton_params* p,
                  dmx_newton_info* info);

/* ---- batched executor -------------------------------------------------------------------
 * Many independent problems (contexts, any mix of dim 2/3 and 1p/2p) advanced together on one
 * GPU with iteration-level continuous batching: each problem's Newton iteration, time-step
 * control and FP64 iterative refinement around FP32 BiCGSTAB + block Jacobi run as a state
 * machine on the device, and one "tick" (a captured HIP graph of one launch per phase for all
 * problems) moves every problem through whatever phases it is in, so problems proceed at their
 * own pace and join or leave between ticks. A problem's result does not depend on the other
 * problems in the batch or on when they join or leave (bitwise).
 * No reference counterpart: the reference runs one problem per process. Per problem the
 * control follows NewtonSolver::solve(vars, timeLoop) (newtonsolver.hh:309-355),
 * newtonProceed/newtonConverged (:428-446, 657-701) and solveImpl_ (:976-1072); the linear
 * solve is the DMX_PRECISION_MIXED + DMX_PRECOND_JACOBI form of dmx_linear_solve
 * (`linear.precond` and `linear.precision` are ignored). */
typedef struct dmx_batch dmx_batch;
typedef struct dmx_batch_item {
    int32_t nsteps;      /* implicit-Euler steps to tEnd = nsteps*dt (ignored for linear_pde) */
    int32_t linear_pde;  /* 1: one stationary LinearPDESolver step */
    double dt;
    double reduction;    /* linear residual reduction of this problem; <= 0: linear.reduction */
} dmx_batch_item;
/* the Newton and linear parameters hold for every problem of the batch */
int dmx_batch_create(int32_t device, const dmx_newton_params* p, dmx_batch** out);
int dmx_batch_destroy(dmx_batch* b);
const char* dmx_batch_last_error(const dmx_batch* b);
/* the context's current solution is the initial (and, instationary, previous) solution; it
 * stays in the context's device arrays and is updated in place */
int dmx_batch_add(dmx_batch* b, dmx_context* ctx, const dmx_batch_item* item);
int dmx_batch_remove(dmx_batch* b, dmx_context* ctx);
int dmx_batch_size(const dmx_batch* b);
/* run `ticks` ticks, then report the number of members still running */
int dmx_batch_step(dmx_batch* b, int32_t ticks, int32_t* nrunning);
/* finished: 1 reached tEnd, -1 failed (dt divisions exhausted), 0 running; as of the last step */
int dmx_batch_status(const dmx_batch* b, dmx_context* ctx, int32_t* finished, dmx_newton_info* info);
/* ticks until no member is running */
int dmx_batch_run(dmx_batch* b);
/* dmx_time_kernel for the batched kernels over all members (DMX_KERNEL_ASSEMBLE,
 * _REFINE_RESIDUAL, _BJ_SPMV_DOT_FP32, _BJ_SPMV_DOTS_FP32, _BJ_ITERATION_FP32); bytes summed
 * over the members. Only when no member is running; their solutions stay untouched. */
int dmx_batch_time_kernel(dmx_batch* b, int32_t kernel, int32_t reps, double* avg_ms, double* bytes);
/* Launch shape of the batched BiCGSTAB kernels: tiles of 256 cells per workgroup pass of the
 * vector kernels (1, 2, 4 or 8), their grid, and the SpMV grid (defaults 4, 2048, 8192).
 * dmx_batch_autotune times `reps` iterations of every candidate (coarse 1/2/4/8 x vector grid
 * 1024..8192 x SpMV grid 2048..16384) on the members' own operators and keeps the fastest;
 * like dmx_batch_time_kernel it needs a step with the current members and none running, and
 * leaves the solutions untouched. The shape changes timings only, never results. */
int dmx_batch_autotune(dmx_batch* b, int32_t reps, double* best_us);
int dmx_batch_set_tuning(dmx_batch* b, int32_t coarse, int32_t vec_grid, int32_t spmv_grid);
int dmx_batch_get_tuning(const dmx_batch* b, int32_t* coarse, int32_t* vec_grid, int32_t* spmv_grid);

/* ---- measurement ------------------------------------------------------------------------
 * Launch one kernel of the path `reps` times back to back on the context's stream between two
 * HIP events and report its average duration and its algorithmic HBM bytes per launch (each
 * input read once, each output written once; structured geometry derived from indices).
 * Requires an assembled system (and, for the Krylov kernels, a previous linear solve that
 * allocated the vectors of the requested precision). No reference counterpart: this is the
 * instrumentation behind the roofline figures. */
enum {
    DMX_KERNEL_ASSEMBLE = 1,      /* residual + FD Jacobian column per cell (FP64) */
    DMX_KERNEL_CELL_STATE = 2,    /* undeflected volume variables (FP64) */
    DMX_KERNEL_RESIDUAL = 3,      /* residual only (FP64) */
    DMX_KERNEL_SPMV_FP64 = 4,     /* BCRS y = A x */
    DMX_KERNEL_SPMV_FP32 = 5,
    DMX_KERNEL_ILU0_APPLY_FP64 = 6, /* both triangular sweeps, all levels */
    DMX_KERNEL_JACOBI_APPLY_FP32 = 7,
    DMX_KERNEL_BICG_HALF_FP32 = 8,  /* x += a y, r -= a v, ||r||^2 */
    DMX_KERNEL_NEWTON_UPDATE = 9,
    DMX_KERNEL_REFINE_RESIDUAL = 10, /* r = b - A x in FP64 */
    DMX_KERNEL_BJ_SPMV_DOT_FP32 = 11,  /* fused v = A y, <rt, v> (block-Jacobi BiCGSTAB) */
    DMX_KERNEL_BJ_SPMV_DOTS_FP32 = 12, /* fused t = A y, <t, r>, <t, t> */
    DMX_KERNEL_BJ_ITERATION_FP32 = 13, /* all five kernels of one fused iteration (with DMX_PRECOND_AMG
                                          also its two V-cycles) */
    DMX_KERNEL_MG_DOWN0 = 14,          /* batch, DMX_PRECOND_AMG: level 0's multigrid down-sweep */
    DMX_KERNEL_MG_UP0 = 15             /* batch, DMX_PRECOND_AMG: level 0's multigrid up-sweep */
};
int dmx_time_kernel(dmx_context* ctx, int32_t kernel, int32_t reps, double* avg_ms, double* bytes);

/* ---- introspection used by tests ------------------------------------------------------- */
/* material constants of the regularized van Genuchten law (24 doubles, host-side init) */
int dmx_material_params(const dmx_problem_desc* desc, double* out);
/* synchronise the context's stream and report the last kernel's device time in ms */
int dmx_synchronize(dmx_context* ctx);

#ifdef __cplusplus
}
#endif
#endif
