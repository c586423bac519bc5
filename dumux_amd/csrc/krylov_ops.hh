// Scalar steps of BiCGSTAB after a reduction (dune-istl BiCGSTABSolver::apply step order,
// third-party): shared by the single-problem finalize kernel (linalg.hip) and the per-problem
// finalize of the batched executor (batch.hip). tot = reduced values of the operation.
#pragma once
#include "kernels.hh"

namespace dmx {

constexpr double kKrylovEps = 1e-80;  // dune-istl's breakdown threshold

__device__ __forceinline__ void fin_apply(int op, const double* tot, KrylovScalars* ks, double* out, double reduction,
                                          int32_t maxit)
{
    switch (op) {
    case FIN_INIT: {
        const double norm = __builtin_sqrt(tot[0]);
        ks->norm = norm;
        ks->def0 = norm;
        ks->reduction_target = reduction;
        ks->rho = 1.0;
        ks->alpha = 1.0;
        ks->omega = 1.0;
        ks->rho_new = tot[0];  // <rt, r> with rt = r
        ks->half_steps = 0;
        ks->status = __builtin_isfinite(norm) ? 0 : 1;
        ks->done = (!__builtin_isfinite(norm) || norm < 1e-30) ? 1 : 0;
        break;
    }
    case FIN_H: {
        const double h = tot[0];
        if (__builtin_fabs(h) < kKrylovEps || !__builtin_isfinite(h)) { ks->status = 1; ks->done = 1; break; }
        ks->h = h;
        ks->alpha = (ks->norm == 0.0) ? 0.0 : ks->rho_new / h;
        break;
    }
    case FIN_HALF1: {
        const double norm = __builtin_sqrt(tot[0]);
        ks->norm = norm;
        ks->half_steps += 1;
        if (!__builtin_isfinite(norm)) { ks->status = 1; ks->done = 1; }
        else if (norm < ks->def0 * ks->reduction_target || norm < 1e-30) ks->done = 1;
        break;
    }
    case FIN_OMEGA:
        ks->tr = tot[0];
        ks->tt = tot[1];
        ks->omega = (ks->norm == 0.0) ? 0.0 : tot[0] / tot[1];
        break;
    case FIN_HALF2: {
        const double norm = __builtin_sqrt(tot[0]);
        ks->norm = norm;
        ks->half_steps += 1;
        ks->rho = ks->rho_new;
        ks->rho_new = tot[1];
        if (!__builtin_isfinite(norm)) { ks->status = 1; ks->done = 1; break; }
        if (norm < ks->def0 * ks->reduction_target || norm < 1e-30) { ks->done = 1; break; }
        // breakdown tests at the top of the next iteration
        if (__builtin_fabs(ks->rho) <= kKrylovEps || __builtin_fabs(ks->omega) <= kKrylovEps) {
            ks->status = 1;
            ks->done = 1;
            break;
        }
        if (ks->half_steps >= 2 * maxit) { ks->status = 1; ks->done = 1; }
        break;
    }
    case FIN_NORM: out[0] = __builtin_sqrt(tot[0]); break;
    case FIN_MAX: out[0] = tot[0]; break;
    case FIN_SUM: out[0] = tot[0]; break;
    default: break;
    }
}

} // namespace dmx
