/* Oracle (test infrastructure only): C API of the CPU restatement of DuMux's CCTpfa 1p/2p
 * Newton path. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it,
 * as the checker. Nothing in the dumux_amd product links or calls it.
 *
 * Conventions (DuMux/YaspGrid, adopted from the reference):
 *   cell index x-fastest; per cell the faces are visited -x,+x,-y,+y(,-z,+z)
 *   (dumux/discretization/cellcentered/tpfa/fvgridgeometry.hh:269-315);
 *   BCRS rows sorted (dumux/assembly/jacobianpattern.hh:25-53); blocks m x m row-major.
 * Boundary faces are enumerated side by side (-x,+x,-y,+y,-z,+z); inside a side by the two
 * tangential cell coordinates, lower axis fastest.
 */
#ifndef DUMUX_ORACLE_H
#define DUMUX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* van Genuchten law of one material (VanGenuchtenDefault parameters) */
typedef struct orc_vg_material {
    double vg_alpha, vg_n, vg_l, swr, snr;
    double pc_low_swe, pc_high_swe, krn_low_swe, krw_high_swe;
} orc_vg_material;

typedef struct orc_problem {
    int32_t dim;          /* 2 or 3 */
    int32_t model;        /* 1: OneP, 2: TwoP (p0s1 = pw-Sn, wetting phase 0) */
    int32_t n[3];         /* cells per axis (n[2] = 1 when dim == 2) */
    int32_t gravity;      /* Problem.EnableGravity: g = -9.81 e_dim */
    double lower[3], upper[3];
    double rho[2], mu[2]; /* phase 0 (wetting/water), phase 1 (DNAPL) */
    double porosity;
    double vg_alpha, vg_n, vg_l, swr, snr;
    double pc_low_swe, pc_high_swe, krn_low_swe, krw_high_swe;
    const double* K;          /* N isotropic permeabilities */
    const int32_t* bc_type;   /* per boundary face: 0 Neumann, 1 Dirichlet */
    const double* bc_value;   /* per boundary face, m values: Dirichlet priVars or Neumann flux per area */
    double fd_base_eps;       /* Assembly.NumericDifference.BaseEpsilon (1e-10) */
    double fd_magnitude[2];   /* Assembly.NumericDifference.PriVarMagnitude (-1: from |x|+1) */
    int32_t diff_method;      /* FVAssembler<TypeTag, DiffMethod>: 0 numeric (forward), 1 analytic */
    /* element-wise materials (fluidMatrixInteraction(element, ...)): material_id[N] in {0, 1}
     * selects the law above (0) or material1 (1); NULL: every cell uses the law above */
    const int32_t* material_id;
    orc_vg_material material1;
} orc_problem;

/* Synthetic problems of the benchmark configurations. kind 1: OneP Darcy, Dirichlet
 * p_left / p_right on -x / +x, no-flow elsewhere. kind 2: TwoP DNAPL infiltration,
 * hydrostatic Dirichlet pw (Sn = 0) on the lateral sides, Neumann DNAPL influx on the central
 * part of the top side, no-flow bottom. */
typedef struct orc_build_spec {
    int32_t kind, dim;
    int32_t n[3];
    uint32_t seed, pid;
    double mu_log10K, sigma_log10K;
    int32_t gravity;
    double p_left, p_right;   /* kind 1 */
    double p_ref;             /* kind 2: pw at the top, also kind-1 initial pressure */
    double influx;            /* kind 2: Neumann DNAPL flux per area (negative = inflow) */
    double inlet_lo, inlet_hi;/* kind 2: inlet = centre within (lo, hi) fractions of each tangential axis */
} orc_build_spec;

typedef struct orc_solve_info {
    int32_t converged;
    int32_t iterations;
    double reduction;
    double def0, def;
} orc_solve_info;

typedef struct orc_newton_params {
    double max_rel_shift;     /* 1e-8 */
    int32_t min_steps;        /* 2 */
    int32_t max_steps;        /* 18 */
    double lin_reduction;     /* LinearSolver.ResidualReduction */
    int32_t lin_max_iter;     /* 250 */
    int32_t precond;          /* 1: ILU0, 2: block Jacobi */
    int32_t max_dt_divisions; /* 10 */
    double retry_factor;      /* 0.5 */
    int32_t linear_pde;       /* 1: one assemble/solve/update (LinearPDESolver) */
    int32_t nthreads;         /* assembly threads (colored) */
    int32_t verbosity;
    int32_t adaptive_dt;      /* 1: dt of the next step = suggestTimeStepSize (time loop only) */
    int32_t target_steps;     /* Newton.TargetSteps 10 */
    double max_dt;            /* TimeLoop max time step size (<= 0: none) */
    int32_t partial_reassembly;      /* Newton.EnablePartialReassembly (numeric Jacobian only) */
    double reassembly_min_threshold; /* Newton.ReassemblyMinThreshold; <= 0: 1e-1 * max_rel_shift */
    double reassembly_max_threshold; /* Newton.ReassemblyMaxThreshold; <= 0: 1e2 * max_rel_shift */
    double reassembly_shift_weight;  /* Newton.ReassemblyShiftWeight 1e-3 */
} orc_newton_params;

typedef struct orc_newton_info {
    int32_t converged;
    int32_t newton_steps;     /* successful Newton iterations summed over time steps */
    int32_t wasted_steps;     /* iterations of failed attempts */
    int32_t linear_iterations;
    int32_t time_steps;
    int32_t dt_retries;
    double last_shift;
    double t_assemble, t_solve, t_update;
    int64_t reassembled_cells;  /* Jacobian columns computed, summed over the assemblies */
} orc_newton_info;

double orc_dpow(double x, double y);
double orc_dlog(double x);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_lognormal_K(uint32_t seed, uint32_t pid, int64_t n, double mu, double sigma, double* K);

/* material law of the problem; out receives 24 doubles (see oracle.cpp) */
void orc_material_params(const orc_problem* P, double* out);
void orc_vg_eval(const orc_problem* P, int64_t n, const double* sw, double* pc, double* krw, double* krn, double* dpc);
/* regularized derivatives dpc/dsw, dkrw/dsw, dkrn/dsw (absolute sw) */
void orc_vg_eval_deriv(const orc_problem* P, int64_t n, const double* sw, double* dpc, double* dkrw, double* dkrn);
/* un-regularized law (absolute sw) */
void orc_vg_eval_noreg(const orc_problem* P, int64_t n, const double* sw, double* pc, double* krw, double* krn);

int64_t orc_num_cells(const orc_problem* P);
int64_t orc_num_boundary_faces(int32_t dim, const int32_t* n);
int orc_build(const orc_build_spec* spec, orc_problem* P, double* K, int32_t* bc_type, double* bc_value, double* u0);

int64_t orc_pattern(const orc_problem* P, int32_t* rowptr, int32_t* colidx);
int orc_coloring(const orc_problem* P, int32_t* colors);
/* r (N*m) and J (nnzb*m*m, may be NULL); uprev == NULL: stationary; J by P->diff_method */
int orc_assemble(const orc_problem* P, const double* u, const double* uprev, double dt,
                 double* r, double* J, int32_t nthreads);

/* PartialReassembler for CCTpfa (partialreassembler.hh:308-346) after a Newton update
 * (newtonsolver.hh:565-603): dist[i] += the maximum relative shift of cell i between uLast and
 * uLast - (uLast - u); green[i] = !(dist[i] > threshold); dist of red cells reset to 0.
 * Returns the number of green cells. */
int64_t orc_partial_colors(int64_t ncells, int32_t m, const double* uLast, const double* u, double threshold,
                           double* dist, uint8_t* green);
/* assembleJacobianAndResidual(u, partialReassembler) (fvassembler.hh:178-207, 434-445,
 * cclocalassembler.hh:61-76): the columns of non-green cells are zeroed and reassembled, green
 * cells only evaluate their residual; J holds the previous Jacobian on entry. Numeric Jacobian. */
int orc_assemble_partial(const orc_problem* P, const double* u, const double* uprev, double dt, const uint8_t* green,
                         double* r, double* J);
int orc_ilu0_factor(int64_t nrows, int32_t m, const int32_t* rowptr, const int32_t* colidx, double* vals);
void orc_ilu0_apply(int64_t nrows, int32_t m, const int32_t* rowptr, const int32_t* colidx,
                    const double* lu, double* v, const double* d);
void orc_spmv(int64_t nrows, int32_t m, const int32_t* rowptr, const int32_t* colidx,
              const double* A, const double* x, double* y);
int orc_bicgstab(int64_t nrows, int32_t m, const int32_t* rowptr, const int32_t* colidx,
                 const double* A, double* x, const double* b, double reduction, int32_t maxit,
                 int32_t precond, orc_solve_info* info);

void orc_default_newton_params(orc_newton_params* p);
/* one Newton solve of one time step (uprev NULL: stationary); returns 0 converged, 1 not */
int orc_newton(const orc_problem* P, double* u, const double* uprev, double dt,
               const orc_newton_params* np, orc_newton_info* info);
/* NewtonSolver::suggestTimeStepSize (newtonsolver.hh:784-798) */
double orc_suggest_time_step(double oldDt, int32_t numSteps, int32_t targetSteps);
/* implicit-Euler loop to t_end = nsteps*dt with dt-halving retries (adaptive_dt: the next dt is
 * the suggested one, capped by max_dt and the end time) */
int orc_time_loop(const orc_problem* P, double* u, int32_t nsteps, double dt,
                  const orc_newton_params* np, orc_newton_info* info);

#ifdef __cplusplus
}
#endif
#endif
