// Launch wrappers of the gfx950 kernels (assembly.hip, linalg.hip). All take the context's
// HIP stream and device pointers; none allocates.
#pragma once
#include "cellops.hh"
#include "mg.hh"
#include <hip/hip_runtime.h>

namespace dmx {

// ---- assembly ----------------------------------------------------------------------------
// vvd (may be null): 2p saturation derivatives of the analytic Jacobian, 3 per cell;
// mat (may be null): material id per cell
hipError_t launch_cell_state(hipStream_t s, const Phys& ph, const double* u, double* vvp, double* vvu, double* vvd,
                             const int32_t* mat, int64_t N);
// bmat (may be null): material of each boundary face's inside cell
hipError_t launch_ghost_state(hipStream_t s, const Phys& ph, const int32_t* bct, const double* bcv, const int32_t* bmat,
                              double* gp, double* gu, int64_t nbf);
hipError_t launch_assemble(hipStream_t s, const GridDesc& g, const Phys& ph, const AsmPtrs& P, bool jacobian);

// ---- block sparse / vector kernels ----------------------------------------------------------
// Device-resident scalars of one Krylov solve (read by the kernels, so no host round trip
// is needed inside an iteration).
struct KrylovScalars {
    double rho, rho_new, alpha, omega, beta, h;
    double norm, def0, reduction_target;
    double tr, tt;
    int32_t done;        // 1: converged, the remaining kernels of the solve are no-ops
    int32_t status;      // 0 running/converged, 1 breakdown or non-finite defect
    int32_t half_steps;  // number of half iterations done (dune-istl counts it += 0.5)
    int32_t pad;
};

struct ReduceWork {
    double* partial;     // 4 * cap
    int64_t cap;         // partial values per slot (>= number of workgroups of any reduction)
};
constexpr int REDUCE_MAX_BLOCKS = 1024;

// Scalar step after a reduction: one workgroup sums the per-workgroup partials (slot k at
// partial[k*cap + b], b < nparts) in index order and updates the Krylov scalars. Producers write
// partials with plain stores; the kernel boundary makes them visible (no device-scope fences,
// which cost an L2 write-back + invalidate per workgroup on gfx950).
enum FinalizeOp {
    FIN_INIT = 0,    // <r,r> of the initial residual
    FIN_H = 1,       // <rt,v>  -> alpha
    FIN_HALF1 = 2,   // <r,r>   -> norm, convergence
    FIN_OMEGA = 3,   // <t,r>, <t,t> -> omega
    FIN_HALF2 = 4,   // <r,r>, <rt,r> -> norm, rho, convergence, breakdown
    FIN_NORM = 5,    // sqrt(sum) -> out[0]
    FIN_MAX = 6,     // max -> out[0]
    FIN_SUM = 7      // sum -> out[0]
};
hipError_t launch_finalize(hipStream_t s, int op, int64_t nparts, KrylovScalars* ks, ReduceWork w, double* out,
                           double reduction, int32_t maxit);

// ---- fused block-Jacobi BiCGSTAB (krylov.hip) ------------------------------------------------
int64_t tile_reduce_blocks(int64_t nrows);
// one BiCGSTAB iteration, vec = {r, rt, p, v, t, y, x}
template <typename T>
hipError_t bj_iteration(hipStream_t s, int m, int64_t nrows, const int32_t* rowptr, const int32_t* colidx,
                        const T* A, const T* Dinv, T* const* vec, KrylovScalars* ks, ReduceWork w, int32_t maxit);
// y = A x with the workgroup's block products staged in LDS (rows of at most 2*dim+1 blocks);
// no-op once ks->done (ks may be null)
template <typename T>
hipError_t launch_spmv_tile(hipStream_t s, int m, int64_t nrows, const int32_t* rowptr, const int32_t* colidx,
                            const T* A, const T* x, T* y, const KrylovScalars* ks = nullptr);
template <typename T>
hipError_t bj_kernel(hipStream_t s, int which, int m, int64_t nrows, const int32_t* rowptr, const int32_t* colidx,
                     const T* A, const T* Dinv, T* const* vec, KrylovScalars* ks, ReduceWork w, int32_t maxit);

// r = b - A x in FP64
hipError_t launch_residual(hipStream_t s, int m, int64_t nrows, const int32_t* rowptr, const int32_t* colidx,
                           const double* A, const double* x, const double* b, double* r);

// ILU(0): level-by-level factorization of LU (a copy of A) and the two triangular sweeps
hipError_t launch_ilu0_factor_level(hipStream_t s, int m, const int32_t* rowptr, const int32_t* colidx,
                                    const int32_t* diag, double* LU, const int32_t* rows, int32_t count,
                                    int32_t* fail);
template <typename T>
hipError_t launch_ilu0_lower_level(hipStream_t s, int m, const int32_t* rowptr, const int32_t* colidx,
                                   const T* LU, const int32_t* rows, int32_t count, const T* d, T* v,
                                   const KrylovScalars* ks);
template <typename T>
hipError_t launch_ilu0_upper_level(hipStream_t s, int m, const int32_t* rowptr, const int32_t* colidx,
                                   const int32_t* diag, const T* LU, const int32_t* rows, int32_t count, T* v,
                                   const KrylovScalars* ks);
// ILU(0) apply y = (LU)^-1 d as two launches (forward, backward) of `grid` workgroups that solve
// boxes of cells drawn in box-level order (`order`: box indices sorted by tx+ty+tz). The pattern
// must be the structured stencil of g. tmp holds the forward sweep's result and must hold the
// sweep pattern on entry (launch_fill_mark once; the backward sweep restores it); tickets
// (2 counters) must be zero before the first apply. fail |= 2 if a box waited past the spin
// limit (the result is then not valid).
template <typename T>
hipError_t launch_ilu0_sweeps(hipStream_t s, int m, const GridDesc& g, const TileGeom& tg, const int32_t* order,
                              const int32_t* rowptr, const int32_t* colidx, const int32_t* diag, const T* LU, int grid,
                              const T* d, T* tmp, T* y, uint32_t* tickets, int32_t* fail, const KrylovScalars* ks);
template <typename T>
hipError_t launch_fill_mark(hipStream_t s, T* x, int64_t n);

// block Jacobi: Dinv = inverse of the diagonal blocks
template <typename T>
hipError_t launch_jacobi_setup(hipStream_t s, int m, int64_t nrows, const int32_t* diag, const double* A, T* Dinv,
                               int32_t* fail);

hipError_t launch_convert_f32(hipStream_t s, const double* x, float* y, int64_t n);
hipError_t launch_copy(hipStream_t s, const double* x, double* y, int64_t n);
hipError_t launch_fill(hipStream_t s, double* x, double v, int64_t n);

// BiCGSTAB building blocks (T = double or float); ks = device scalars
template <typename T>
hipError_t bicg_init(hipStream_t s, int64_t n, const double* b, T* r, T* rt, T* p, T* v, T* x,
                     KrylovScalars* ks, ReduceWork w, double reduction, double bscale);
template <typename T>
hipError_t bicg_update_p(hipStream_t s, int64_t n, const T* r, const T* v, T* p, int first, KrylovScalars* ks);
template <typename T>
hipError_t bicg_half1(hipStream_t s, int64_t n, const T* y, const T* v, T* x, T* r, KrylovScalars* ks, ReduceWork w);
template <typename T>
hipError_t bicg_half2(hipStream_t s, int64_t n, const T* y, const T* t, const T* rt, T* x, T* r,
                      KrylovScalars* ks, ReduceWork w, int32_t maxit);
template <typename T>
hipError_t launch_jacobi_apply(hipStream_t s, int m, int64_t nrows, const T* Dinv, const T* p, T* y,
                               const KrylovScalars* ks);

// Newton update u = uLast - du and max relative shift (device max reduction)
hipError_t launch_newton_update(hipStream_t s, int64_t n, const double* uLast, const double* du, double* u,
                                double* shift_out, ReduceWork w);
// ||x||_2 (sequential-order-independent deterministic two-level reduction)
hipError_t launch_norm(hipStream_t s, int64_t n, const double* x, double* out, ReduceWork w);
hipError_t launch_partial_colors(hipStream_t s, int64_t ncells, int m, const double* uLast, const double* u, double threshold,
                                 double* dist, uint8_t* green, double* ngreen_out, ReduceWork w);
// x += scale*y (FP64) with y in T
template <typename T>
hipError_t launch_add_to(hipStream_t s, int64_t n, const T* y, double scale, double* x);
// r32 = float(r)
hipError_t launch_to_f32(hipStream_t s, int64_t n, const double* r, float* r32);

// ---- multigrid preconditioner of the single-problem BiCGSTAB (multigrid.hip) ----------------
// the levels' operators from the FP64 BCRS Jacobian (level 0 unscaled, explicit diagonal)
hipError_t launch_mg1_setup(hipStream_t s, int m, const GridDesc& g, const int32_t* rowptr, const double* J,
                            const MgLevel* lev, int nl, int32_t* fail);
// out = M^-1 in (one FP32 V-cycle; FP64 vectors are rounded on load and widened on store);
// no-op once ks->done. lev: the levels on the host, levDev: the same in device memory;
// lt = mg_tail_level.
template <typename T>
hipError_t launch_mg1_apply(hipStream_t s, int m, const MgLevel* lev, const MgLevel* levDev, int nl, int lt, const T* in,
                            T* out, const KrylovScalars* ks);

} // namespace dmx
