// C ABI of dumux_amd (include/dumux_amd.h): one context = one problem on one HIP stream.
//
// Device memory per context (N cells, m unknowns per cell, nnzb = BCRS blocks):
//   pattern   rowptr (N+1), colidx (nnzb), diag (N), ILU level rows (N)         int32
//   problem   K (N), boundary type/value, Dirichlet ghost states                 FP64
//   state     u, u_last, uprev (N*m each), cell states vv_p / vv_up (N*m)        FP64
//   system    J (nnzb*m*m), r, du (N*m)                                          FP64
//   solver    ILU factors LU (nnzb*m*m), block-Jacobi inverse (N*m*m), 7 Krylov
//             vectors in FP64 and/or FP32, FP32 copy of J for mixed precision (lazy)
// Host-side control (Newton loop, time loop, retries) follows dumux/nonlinear/newtonsolver.hh.
#include "context.hh"
#include "mg.hh"
#include "../../include/dumux_amd/gpu.hh"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace dmx;

namespace {

void free_all(dmx_context* ctx)
{
    auto f = [](void* p) { if (p) (void)hipFree(p); };
    f(ctx->rowptr); f(ctx->colidx); f(ctx->diag); f(ctx->level_rows); f(ctx->bct);
    f(ctx->K); f(ctx->bcv); f(ctx->ghost_p); f(ctx->ghost_up);
    f(ctx->u); f(ctx->u_last); f(ctx->uprev); f(ctx->vvp); f(ctx->vvu); f(ctx->vvd); f(ctx->mat); f(ctx->bmat);
    f(ctx->J); f(ctx->r); f(ctx->du); f(ctx->LU); f(ctx->Dinv64);
    f(ctx->J32); f(ctx->Dinv32); f(ctx->LU32); f(ctx->S32); f(ctx->kvb); f(ctx->mgBuf);
    f(ctx->prDist); f(ctx->prGreen); f(ctx->mg1Buf); f(ctx->mg1LevDev);
    for (int i = 0; i < 7; ++i) { f(ctx->kv64[i]); f(ctx->kv32[i]); }
    f(ctx->ir_r); f(ctx->ir_x); f(ctx->ks); f(ctx->scal); f(ctx->fail);
    f(ctx->sweepTickets); f(ctx->sweepTmp64); f(ctx->sweepTmp32); f(ctx->boxOrder);
    f(ctx->rw.partial);
    for (int i = 0; i < 2; ++i) {
        if (ctx->bjExec[i]) (void)hipGraphExecDestroy(ctx->bjExec[i]);
        if (ctx->bjGraph[i]) (void)hipGraphDestroy(ctx->bjGraph[i]);
    }
    if (ctx->ks_host) (void)hipHostFree(ctx->ks_host);
    if (ctx->scal_host) (void)hipHostFree(ctx->scal_host);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
}

int validate(const dmx_problem_desc* d, std::string& err)
{
    if (!d) { err = "null problem"; return DMX_ERR_INVALID; }
    if (d->dim < 2 || d->dim > 3) { err = "dim must be 2 or 3"; return DMX_ERR_INVALID; }
    if (d->model != DMX_MODEL_ONEP && d->model != DMX_MODEL_TWOP) { err = "unknown model"; return DMX_ERR_INVALID; }
    int64_t N = 1;
    for (int a = 0; a < d->dim; ++a) {
        if (d->n[a] < 1) { err = "cells per axis must be >= 1"; return DMX_ERR_INVALID; }
        if (!(d->upper[a] > d->lower[a])) { err = "empty box"; return DMX_ERR_INVALID; }
        N *= d->n[a];
    }
    if (N > (int64_t(1) << 30) / 8) { err = "grid too large for int32 BCRS indices"; return DMX_ERR_INVALID; }
    if (!d->K || !d->bc_type || !d->bc_value) { err = "K / boundary arrays missing"; return DMX_ERR_INVALID; }
    if (!(d->fd_base_eps > 0.0)) { err = "fd_base_eps must be positive"; return DMX_ERR_INVALID; }
    if (d->material_id)
        for (int64_t i = 0; i < N; ++i)
            if (d->material_id[i] != 0 && d->material_id[i] != 1) {
                err = "material_id must be 0 or 1";
                return DMX_ERR_INVALID;
            }
    if (d->diff_method != DMX_DIFF_NUMERIC && d->diff_method != DMX_DIFF_ANALYTIC) {
        err = "diff_method must be DMX_DIFF_NUMERIC or DMX_DIFF_ANALYTIC";
        return DMX_ERR_INVALID;
    }
    return DMX_OK;
}

AsmPtrs asm_ptrs(dmx_context* ctx, bool jac)
{
    AsmPtrs P{};
    P.u = ctx->u;
    P.uprev = ctx->havePrev ? ctx->uprev : nullptr;
    P.K = ctx->K;
    P.bc_type = ctx->bct;
    P.bc_value = ctx->bcv;
    P.ghost_p = ctx->ghost_p;
    P.ghost_up = ctx->ghost_up;
    P.vv_p = ctx->vvp;
    P.vv_up = ctx->vvu;
    P.vv_d = ctx->vvd;
    P.mat = ctx->mat;
    P.rowptr = ctx->rowptr;
    P.r = ctx->r;
    P.J = jac ? ctx->J : nullptr;
    P.S32 = nullptr;
    P.dt = ctx->dt;
    P.rowBase = 0;
    P.analytic = ctx->desc.diff_method == DMX_DIFF_ANALYTIC ? 1 : 0;
    return P;
}

int ensure_krylov(dmx_context* ctx, bool f32)
{
    if (f32) {
        for (int i = 0; i < 7; ++i)
            if (!ctx->kv32[i]) DMX_TRY(dalloc(ctx, &ctx->kv32[i], ctx->n));
    } else {
        for (int i = 0; i < 7; ++i)
            if (!ctx->kv64[i]) DMX_TRY(dalloc(ctx, &ctx->kv64[i], ctx->n));
    }
    return DMX_OK;
}

// Preconditioner setup on the current J. ILU(0): LU = J, then factor level by level.
// the single-problem multigrid hierarchy (allocated once per context): level 0 holds the FP32
// stencil operator of J (S, D, D^-1) and the pre-smoothed iterate, the coarse levels S, D, D^-1,
// x, f (mg.hh)
int mg1_prepare(dmx_context* ctx)
{
    if (!ctx->mgLev.empty()) return DMX_OK;
    const int nl = mg_num_levels(ctx->g);
    if (nl < 2 || !ctx->sweepFits) {
        ctx->err = "multigrid: the grid has no coarse levels (or the pattern is not the structured stencil)";
        return DMX_ERR_UNSUPPORTED;
    }
    // buffers of an earlier attempt that failed part way
    for (void* q : {static_cast<void*>(ctx->mg1Buf), static_cast<void*>(ctx->mg1LevDev)})
        if (q) (void)hipFree(q);
    ctx->mg1Buf = nullptr;
    ctx->mg1LevDev = nullptr;
    const int64_t m = ctx->m, mm = m * m;
    int64_t total = (int64_t(2) * ctx->g.dim * mm + 2 * mm + m) * ctx->g.N;
    {
        GridDesc c = ctx->g;
        for (int l = 1; l < nl; ++l) {
            c = mg_coarse_grid(c);
            total += (int64_t(2) * c.dim * mm + 2 * mm + 2 * m) * c.N;
        }
    }
    DMX_TRY(dalloc(ctx, &ctx->mg1Buf, total));
    float* q = ctx->mg1Buf;
    std::vector<MgLevel> lev(nl);
    for (int l = 0; l < nl; ++l) {
        const GridDesc g = l == 0 ? ctx->g : mg_coarse_grid(lev[l - 1].g);
        MgLevel& L = lev[l];
        L = MgLevel{};
        L.g = g;
        L.box = mg_boxes(g);
        L.S = q; q += int64_t(2) * g.dim * mm * g.N;
        L.D = q; q += mm * g.N;
        L.Dinv = q; q += mm * g.N;
        L.x = q; q += m * g.N;
        if (l > 0) { L.f = q; q += m * g.N; }
    }
    DMX_TRY(dalloc(ctx, &ctx->mg1LevDev, nl));
    DMX_HIP(hipMemcpy(ctx->mg1LevDev, lev.data(), sizeof(MgLevel) * lev.size(), hipMemcpyHostToDevice));
    ctx->mg1Tail = mg_tail_level(ctx->g, nl);
    ctx->mgLev = lev;
    return DMX_OK;
}

int precond_setup(dmx_context* ctx, int precond, bool f32)
{
    const int mm = ctx->m * ctx->m;
    DMX_HIP(hipMemsetAsync(ctx->fail, 0, sizeof(int32_t), ctx->stream));
    if (precond == DMX_PRECOND_ILU0) {
        if (!ctx->LU) DMX_TRY(dalloc(ctx, &ctx->LU, ctx->nnzb * mm));
        DMX_HIP(hipMemcpyAsync(ctx->LU, ctx->J, sizeof(double) * ctx->nnzb * mm, hipMemcpyDeviceToDevice, ctx->stream));
        for (int l = 0; l < ctx->nlev; ++l) {
            const int32_t beg = ctx->h_level_ptr[l], cnt = ctx->h_level_ptr[l + 1] - beg;
            DMX_HIP(launch_ilu0_factor_level(ctx->stream, ctx->m, ctx->rowptr, ctx->colidx, ctx->diag, ctx->LU,
                                             ctx->level_rows + beg, cnt, ctx->fail));
        }
        if (f32) {
            if (!ctx->LU32) DMX_TRY(dalloc(ctx, &ctx->LU32, ctx->nnzb * mm));
            DMX_HIP(launch_convert_f32(ctx->stream, ctx->LU, ctx->LU32, ctx->nnzb * mm));
        }
        if (ctx->sweepFits && ctx->iluSweep == 0) {
            if (!ctx->sweepTickets) DMX_TRY(dalloc(ctx, &ctx->sweepTickets, 2));
            DMX_HIP(hipMemsetAsync(ctx->sweepTickets, 0, 2 * sizeof(uint32_t), ctx->stream));
            if (f32) {
                if (!ctx->sweepTmp32) DMX_TRY(dalloc(ctx, &ctx->sweepTmp32, ctx->n));
                DMX_HIP(launch_fill_mark<float>(ctx->stream, ctx->sweepTmp32, ctx->n));
            } else {
                if (!ctx->sweepTmp64) DMX_TRY(dalloc(ctx, &ctx->sweepTmp64, ctx->n));
                DMX_HIP(launch_fill_mark<double>(ctx->stream, ctx->sweepTmp64, ctx->n));
            }
        }
    } else if (precond == DMX_PRECOND_AMG) {
        DMX_TRY(mg1_prepare(ctx));
        DMX_HIP(launch_mg1_setup(ctx->stream, ctx->m, ctx->g, ctx->rowptr, ctx->J, ctx->mgLev.data(), int(ctx->mgLev.size()),
                                 ctx->fail));
    } else if (precond == DMX_PRECOND_JACOBI) {
        if (f32) {
            if (!ctx->Dinv32) DMX_TRY(dalloc(ctx, &ctx->Dinv32, ctx->N * mm));
            DMX_HIP(launch_jacobi_setup<float>(ctx->stream, ctx->m, ctx->N, ctx->diag, ctx->J, ctx->Dinv32, ctx->fail));
        } else {
            if (!ctx->Dinv64) DMX_TRY(dalloc(ctx, &ctx->Dinv64, ctx->N * mm));
            DMX_HIP(launch_jacobi_setup<double>(ctx->stream, ctx->m, ctx->N, ctx->diag, ctx->J, ctx->Dinv64, ctx->fail));
        }
    } else {
        ctx->err = "unknown preconditioner";
        return DMX_ERR_INVALID;
    }
    int32_t failed = 0;
    DMX_HIP(hipMemcpyAsync(&failed, ctx->fail, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    if (failed) {
        ctx->err = "singular pivot block in the preconditioner";
        return DMX_NUMERICAL_PROBLEM;
    }
    return DMX_OK;
}

template <typename T>
int precond_apply(dmx_context* ctx, int precond, const T* LUt, const T* Dinv, const T* p, T* y)
{
    if (precond == DMX_PRECOND_ILU0 && ctx->sweepFits && ctx->iluSweep == 0) {
        T* tmp;
        if constexpr (sizeof(T) == 8) tmp = ctx->sweepTmp64;
        else tmp = ctx->sweepTmp32;
        DMX_HIP(launch_ilu0_sweeps<T>(ctx->stream, ctx->m, ctx->g, ctx->tg, ctx->boxOrder, ctx->rowptr, ctx->colidx,
                                      ctx->diag, LUt, ctx->sweepGrid, p, tmp, y, ctx->sweepTickets, ctx->fail, ctx->ks));
    } else if (precond == DMX_PRECOND_ILU0) {
        for (int l = 0; l < ctx->nlev; ++l) {
            const int32_t beg = ctx->h_level_ptr[l], cnt = ctx->h_level_ptr[l + 1] - beg;
            DMX_HIP(launch_ilu0_lower_level<T>(ctx->stream, ctx->m, ctx->rowptr, ctx->colidx, LUt, ctx->level_rows + beg,
                                               cnt, p, y, ctx->ks));
        }
        for (int l = ctx->nlev - 1; l >= 0; --l) {
            const int32_t beg = ctx->h_level_ptr[l], cnt = ctx->h_level_ptr[l + 1] - beg;
            DMX_HIP(launch_ilu0_upper_level<T>(ctx->stream, ctx->m, ctx->rowptr, ctx->colidx, ctx->diag, LUt,
                                               ctx->level_rows + beg, cnt, y, ctx->ks));
        }
    } else if (precond == DMX_PRECOND_AMG) {
        DMX_HIP(launch_mg1_apply<T>(ctx->stream, ctx->m, ctx->mgLev.data(), ctx->mg1LevDev, int(ctx->mgLev.size()),
                                    ctx->mg1Tail, p, y, ctx->ks));
    } else {
        DMX_HIP(launch_jacobi_apply<T>(ctx->stream, ctx->m, ctx->N, Dinv, p, y, ctx->ks));
    }
    return DMX_OK;
}

constexpr int kGraphIters = 8;

bool graphs_enabled()
{
    const char* e = std::getenv("DMX_NO_GRAPH");
    return !(e && e[0] == '1');
}

// BiCGSTAB with block Jacobi: five fused kernels per iteration, replayed as a captured graph of
// kGraphIters iterations; the host reads the device flag once per replay.
template <typename T>
int bicgstab_bj_fused(dmx_context* ctx, const T* A, const T* Dinv, T* const* vec, int maxit)
{
    hipStream_t s = ctx->stream;
    const int slot = sizeof(T) == 4 ? 0 : 1;
    const bool useGraph = graphs_enabled();
    if (useGraph && (!ctx->bjExec[slot] || ctx->bjMaxit[slot] != maxit)) {
        if (ctx->bjExec[slot]) { (void)hipGraphExecDestroy(ctx->bjExec[slot]); ctx->bjExec[slot] = nullptr; }
        if (ctx->bjGraph[slot]) { (void)hipGraphDestroy(ctx->bjGraph[slot]); ctx->bjGraph[slot] = nullptr; }
        DMX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int g = 0; g < kGraphIters; ++g) {
            hipError_t e = bj_iteration<T>(s, ctx->m, ctx->N, ctx->rowptr, ctx->colidx, A, Dinv, vec, ctx->ks, ctx->rw, maxit);
            if (e != hipSuccess) {
                hipGraph_t tmp;
                (void)hipStreamEndCapture(s, &tmp);
                ctx->err = std::string("graph capture: ") + hipGetErrorString(e);
                return DMX_ERR_DEVICE;
            }
        }
        DMX_HIP(hipStreamEndCapture(s, &ctx->bjGraph[slot]));
        DMX_HIP(hipGraphInstantiate(&ctx->bjExec[slot], ctx->bjGraph[slot], nullptr, nullptr, 0));
        ctx->bjMaxit[slot] = maxit;
    }
    for (int it = 0; it < maxit; it += kGraphIters) {
        if (useGraph) {
            DMX_HIP(hipGraphLaunch(ctx->bjExec[slot], s));
        } else {
            for (int g = 0; g < kGraphIters; ++g)
                DMX_HIP(bj_iteration<T>(s, ctx->m, ctx->N, ctx->rowptr, ctx->colidx, A, Dinv, vec, ctx->ks, ctx->rw, maxit));
        }
        DMX_HIP(hipMemcpyAsync(ctx->ks_host, ctx->ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, s));
        DMX_HIP(hipStreamSynchronize(s));
        if (ctx->ks_host->done) break;
    }
    return DMX_OK;
}

// One preconditioned BiCGSTAB iteration (two half steps); vec = {r, rt, p, v, t, y, x}. Each
// SpMV runs fused with the dot products of its result (the block-Jacobi path's kernels B and D).
template <typename T>
int bicg_iteration(dmx_context* ctx, int precond, const T* A, const T* LUt, const T* Dinv, T* const* vec, int maxit,
                   bool first)
{
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n;
    T *r = vec[0], *rt = vec[1], *p = vec[2], *v = vec[3], *t = vec[4], *y = vec[5], *x = vec[6];
    DMX_HIP(bicg_update_p<T>(s, n, r, v, p, first ? 1 : 0, ctx->ks));
    DMX_TRY(precond_apply<T>(ctx, precond, LUt, Dinv, p, y));
    const int64_t np = tile_reduce_blocks(ctx->N);
    DMX_HIP(bj_kernel<T>(s, 1, ctx->m, ctx->N, ctx->rowptr, ctx->colidx, A, Dinv, vec, ctx->ks, ctx->rw, maxit));
    DMX_HIP(launch_finalize(s, FIN_H, np, ctx->ks, ctx->rw, nullptr, 0.0, 0));
    DMX_HIP(bicg_half1<T>(s, n, y, v, x, r, ctx->ks, ctx->rw));
    DMX_TRY(precond_apply<T>(ctx, precond, LUt, Dinv, r, y));
    DMX_HIP(bj_kernel<T>(s, 3, ctx->m, ctx->N, ctx->rowptr, ctx->colidx, A, Dinv, vec, ctx->ks, ctx->rw, maxit));
    DMX_HIP(launch_finalize(s, FIN_OMEGA, np, ctx->ks, ctx->rw, nullptr, 0.0, 0));
    DMX_HIP(bicg_half2<T>(s, n, y, t, rt, x, r, ctx->ks, ctx->rw, maxit));
    return DMX_OK;
}

// BiCGSTAB on A (type T values) for right-hand side b (FP64), solution into x (type T).
// Returns DMX_OK with *converged set; iterations counted in full steps.
template <typename T>
int bicgstab(dmx_context* ctx, const T* A, const T* LUt, const T* Dinv, int precond, const double* b, double bscale,
             T* x, double reduction, int maxit, int* converged, int* iterations)
{
    T *r = nullptr, *rt, *p, *v, *t, *y;
    if constexpr (sizeof(T) == 8) {
        r = reinterpret_cast<T*>(ctx->kv64[0]); rt = reinterpret_cast<T*>(ctx->kv64[1]); p = reinterpret_cast<T*>(ctx->kv64[2]);
        v = reinterpret_cast<T*>(ctx->kv64[3]); t = reinterpret_cast<T*>(ctx->kv64[4]); y = reinterpret_cast<T*>(ctx->kv64[5]);
    } else {
        r = reinterpret_cast<T*>(ctx->kv32[0]); rt = reinterpret_cast<T*>(ctx->kv32[1]); p = reinterpret_cast<T*>(ctx->kv32[2]);
        v = reinterpret_cast<T*>(ctx->kv32[3]); t = reinterpret_cast<T*>(ctx->kv32[4]); y = reinterpret_cast<T*>(ctx->kv32[5]);
    }
    const int64_t n = ctx->n;
    hipStream_t s = ctx->stream;
    DMX_HIP(bicg_init<T>(s, n, b, r, rt, p, v, x, ctx->ks, ctx->rw, reduction, bscale));
    if (precond == DMX_PRECOND_JACOBI) {
        T* vec[7] = {r, rt, p, v, t, y, x};
        DMX_TRY(bicgstab_bj_fused<T>(ctx, A, Dinv, vec, maxit));
        const KrylovScalars& k = *ctx->ks_host;
        *converged = (k.done && k.status == 0) ? 1 : 0;
        *iterations = (k.half_steps + 1) / 2;
        return DMX_OK;
    }
    T* vec[7] = {r, rt, p, v, t, y, x};
    // poll the device flag every iteration for ILU(0) (one iteration is hundreds of launches)
    const int checkEvery = (precond == DMX_PRECOND_ILU0 && !(ctx->sweepFits && ctx->iluSweep == 0)) ? 1 : 4;
    int it = 0;
    for (; it < maxit; ++it) {
        DMX_TRY(bicg_iteration<T>(ctx, precond, A, LUt, Dinv, vec, maxit, it == 0));
        if ((it + 1) % checkEvery == 0 || it + 1 == maxit) {
            DMX_HIP(hipMemcpyAsync(ctx->ks_host, ctx->ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, s));
            DMX_HIP(hipStreamSynchronize(s));
            if (ctx->ks_host->done) break;
        }
    }
    DMX_HIP(hipMemcpyAsync(ctx->ks_host, ctx->ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, s));
    DMX_HIP(hipStreamSynchronize(s));
    if (precond == DMX_PRECOND_ILU0) {
        int32_t failed = 0;
        DMX_HIP(hipMemcpy(&failed, ctx->fail, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (failed & 2) {
            ctx->err = "ILU(0) sweep: a row waited past the spin limit";
            return DMX_ERR_DEVICE;
        }
    }
    const KrylovScalars& k = *ctx->ks_host;
    *converged = (k.done && k.status == 0) ? 1 : 0;
    *iterations = (k.half_steps + 1) / 2;
    return DMX_OK;
}

int device_norm(dmx_context* ctx, const double* x, double* out)
{
    DMX_HIP(launch_norm(ctx->stream, ctx->n, x, ctx->scal + 1, ctx->rw));
    DMX_HIP(hipMemcpyAsync(ctx->scal_host + 1, ctx->scal + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    *out = ctx->scal_host[1];
    return DMX_OK;
}

} // namespace

extern "C" {

int64_t dmx_num_boundary_faces(int32_t dim, const int32_t* n)
{
    if (dim < 2 || dim > 3 || !n) return -1;
    int64_t N = 1;
    for (int a = 0; a < dim; ++a) N *= n[a];
    int64_t s = 0;
    for (int a = 0; a < dim; ++a) s += 2 * (N / n[a]);
    return s;
}

int dmx_build_problem(const dmx_build_spec* spec, dmx_problem_desc* desc, double* K, int32_t* bc_type,
                      double* bc_value, double* u0)
{
    if (!spec || !desc || !K || !bc_type || !bc_value || !u0) return DMX_ERR_INVALID;
    return build_problem(*spec, *desc, K, bc_type, bc_value, u0);
}

int dmx_lognormal_permeability(uint32_t seed, uint32_t pid, int64_t n, double mu, double sigma, double* K)
{
    if (!K || n < 0) return DMX_ERR_INVALID;
    for (int64_t c = 0; c < n; ++c) K[c] = lognormal_permeability(seed, pid, uint32_t(c), mu, sigma);
    return DMX_OK;
}

int dmx_material_params(const dmx_problem_desc* d, double* out)
{
    if (!d || !out) return DMX_ERR_INVALID;
    VGLaw L;
    vg_init(L, d->vg_alpha, d->vg_n, d->vg_l, d->swr, d->snr, d->pc_low_swe, d->pc_high_swe, d->krn_low_swe,
            d->krw_high_swe);
    const double v[24] = {L.alpha, L.n, L.m, L.l, L.swr, L.snr,
                          L.pcLowSwe, L.pcHighSwe, L.krwHighSwe, L.krnLowSwe,
                          L.pcLowSwePcValue, L.pcHighSwePcValue, L.pcDerivativeLowSw,
                          L.pcDerivativeHighSweThreshold, L.pcDerivativeHighSweEnd,
                          L.pcSpline.m0, L.pcSpline.m1,
                          L.krwSpline.y0, L.krwSpline.m0, L.krwSpline.m1,
                          L.krnSpline.y1, L.krnSpline.m0, L.krnSpline.m1, 0.0};
    for (int i = 0; i < 24; ++i) out[i] = v[i];
    return DMX_OK;
}

int dmx_create(const dmx_problem_desc* desc, int32_t device, dmx_context** out)
{
    if (!out) return DMX_ERR_INVALID;
    *out = nullptr;
    std::string err;
    const int vr = validate(desc, err);
    if (vr != DMX_OK) {
        std::fprintf(stderr, "dmx_create: %s\n", err.c_str());
        return vr;
    }
    auto* ctx = new dmx_context();
    ctx->device = device;
    ctx->desc = *desc;
    ctx->desc.K = nullptr;
    ctx->desc.bc_type = nullptr;
    ctx->desc.bc_value = nullptr;
    ctx->desc.material_id = nullptr;
    ctx->g = make_grid(*desc);
    ctx->ph = make_phys(*desc);
    ctx->m = ctx->ph.m;
    ctx->N = ctx->g.N;
    ctx->n = ctx->N * ctx->m;
    ctx->nbf = num_boundary_faces(ctx->g);
    build_pattern(ctx->g, ctx->h_rowptr, ctx->h_colidx);
    ctx->nnzb = int64_t(ctx->h_colidx.size());
    ctx->h_diag.resize(ctx->N);
    for (int64_t i = 0; i < ctx->N; ++i)
        for (int32_t k = ctx->h_rowptr[i]; k < ctx->h_rowptr[i + 1]; ++k)
            if (ctx->h_colidx[k] == i) ctx->h_diag[i] = k;
    ctx->nlev = level_schedule(ctx->N, ctx->h_rowptr.data(), ctx->h_colidx.data(), ctx->h_level_ptr, ctx->h_level_rows);
    ctx->sweepFits = structured_stencil_pattern(ctx->g, ctx->h_rowptr, ctx->h_colidx);
    if (ctx->sweepFits) {
        ctx->tg = sweep_boxes(ctx->g, ctx->h_boxOrder);
        ctx->sweepGrid = ctx->g.dim == 3 ? 512 : 256;   // workgroups (box-sized) per sweep launch
    }

    auto fail = [&](int rc) {
        std::fprintf(stderr, "dmx_create: %s\n", ctx->err.c_str());
        free_all(ctx);
        delete ctx;
        return rc;
    };
#define DMX_CREATE_HIP(call)                                                        \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);           \
            return fail(DMX_ERR_DEVICE);                                            \
        }                                                                           \
    } while (0)
#define DMX_CREATE_TRY(expr)              \
    do {                                  \
        int rc_ = (expr);                 \
        if (rc_ != DMX_OK) return fail(rc_); \
    } while (0)

    DMX_CREATE_HIP(hipSetDevice(device));
    DMX_CREATE_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    const int m = ctx->m, mm = m * m;
    DMX_CREATE_TRY(dalloc(ctx, &ctx->rowptr, ctx->N + 1));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->colidx, ctx->nnzb));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->diag, ctx->N));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->level_rows, ctx->N));
    if (ctx->sweepFits) DMX_CREATE_TRY(dalloc(ctx, &ctx->boxOrder, ctx->tg.ntiles));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->bct, ctx->nbf));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->K, ctx->N));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->bcv, ctx->nbf * m));
    std::vector<int32_t> h_bmat;
    if (desc->material_id) {
        DMX_CREATE_TRY(dalloc(ctx, &ctx->mat, ctx->N));
        DMX_CREATE_TRY(dalloc(ctx, &ctx->bmat, ctx->nbf));
        boundary_materials(ctx->g, desc->material_id, h_bmat);
    }
    DMX_CREATE_TRY(dalloc(ctx, &ctx->ghost_p, ctx->nbf * m));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->ghost_up, ctx->nbf * m));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->u, ctx->n));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->u_last, ctx->n));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->uprev, ctx->n));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->vvp, ctx->n));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->vvu, ctx->n));
    if (ctx->m == 2 && desc->diff_method == DMX_DIFF_ANALYTIC) DMX_CREATE_TRY(dalloc(ctx, &ctx->vvd, 3 * ctx->N));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->J, ctx->nnzb * mm));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->r, ctx->n));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->du, ctx->n));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->ks, 1));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->scal, 8));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->fail, 1));
    ctx->rw.cap = std::max<int64_t>(REDUCE_MAX_BLOCKS, tile_reduce_blocks(ctx->N));
    DMX_CREATE_TRY(dalloc(ctx, &ctx->rw.partial, 4 * ctx->rw.cap));
    DMX_CREATE_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->ks_host), sizeof(KrylovScalars), hipHostMallocDefault));
    DMX_CREATE_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->scal_host), 8 * sizeof(double), hipHostMallocDefault));
    hipStream_t s = ctx->stream;
    DMX_CREATE_HIP(hipMemsetAsync(ctx->ks, 0, sizeof(KrylovScalars), s));
    DMX_CREATE_HIP(hipMemcpyAsync(ctx->rowptr, ctx->h_rowptr.data(), sizeof(int32_t) * (ctx->N + 1), hipMemcpyHostToDevice, s));
    DMX_CREATE_HIP(hipMemcpyAsync(ctx->colidx, ctx->h_colidx.data(), sizeof(int32_t) * ctx->nnzb, hipMemcpyHostToDevice, s));
    DMX_CREATE_HIP(hipMemcpyAsync(ctx->diag, ctx->h_diag.data(), sizeof(int32_t) * ctx->N, hipMemcpyHostToDevice, s));
    DMX_CREATE_HIP(hipMemcpyAsync(ctx->level_rows, ctx->h_level_rows.data(), sizeof(int32_t) * ctx->N, hipMemcpyHostToDevice, s));
    if (ctx->sweepFits)
        DMX_CREATE_HIP(hipMemcpyAsync(ctx->boxOrder, ctx->h_boxOrder.data(), sizeof(int32_t) * ctx->tg.ntiles,
                                      hipMemcpyHostToDevice, s));
    DMX_CREATE_HIP(hipMemcpyAsync(ctx->bct, desc->bc_type, sizeof(int32_t) * ctx->nbf, hipMemcpyHostToDevice, s));
    DMX_CREATE_HIP(hipMemcpyAsync(ctx->K, desc->K, sizeof(double) * ctx->N, hipMemcpyHostToDevice, s));
    DMX_CREATE_HIP(hipMemcpyAsync(ctx->bcv, desc->bc_value, sizeof(double) * ctx->nbf * m, hipMemcpyHostToDevice, s));
    if (desc->material_id) {
        DMX_CREATE_HIP(hipMemcpyAsync(ctx->mat, desc->material_id, sizeof(int32_t) * ctx->N, hipMemcpyHostToDevice, s));
        DMX_CREATE_HIP(hipMemcpyAsync(ctx->bmat, h_bmat.data(), sizeof(int32_t) * ctx->nbf, hipMemcpyHostToDevice, s));
    }
    DMX_CREATE_HIP(launch_ghost_state(s, ctx->ph, ctx->bct, ctx->bcv, ctx->bmat, ctx->ghost_p, ctx->ghost_up, ctx->nbf));
    DMX_CREATE_HIP(hipStreamSynchronize(s));
#undef DMX_CREATE_HIP
#undef DMX_CREATE_TRY
    *out = ctx;
    return DMX_OK;
}

int dmx_destroy(dmx_context* ctx)
{
    if (!ctx) return DMX_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
    return DMX_OK;
}

const char* dmx_last_error(const dmx_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int dmx_device_bytes(const dmx_context* ctx, int64_t* bytes)
{
    if (!ctx || !bytes) return DMX_ERR_INVALID;
    *bytes = ctx->devBytes;
    return DMX_OK;
}

int dmx_estimate_bytes(const dmx_problem_desc* desc, int32_t batched, int64_t* bytes)
{
    if (!desc || !bytes || desc->dim < 2 || desc->dim > 3) return DMX_ERR_INVALID;
    if (desc->model != DMX_MODEL_ONEP && desc->model != DMX_MODEL_TWOP) return DMX_ERR_INVALID;
    for (int d = 0; d < desc->dim; ++d)
        if (desc->n[d] < 1) return DMX_ERR_INVALID;
    const GridDesc g = make_grid(*desc);
    const int64_t m = desc->model == DMX_MODEL_TWOP ? 2 : 1, mm = m * m;
    const int64_t N = g.N, n = N * m, nbf = num_boundary_faces(g);
    // structured stencil: the diagonal plus both directions of every interior face
    int64_t nnzb = N;
    for (int d = 0; d < g.dim; ++d) {
        int64_t faces = g.n[d] - 1;
        for (int e = 0; e < g.dim; ++e)
            if (e != d) faces *= g.n[e];
        nnzb += 2 * faces;
    }
    const int64_t cap = std::max<int64_t>(REDUCE_MAX_BLOCKS, tile_reduce_blocks(N));
    // the allocations of dmx_create
    std::vector<int32_t> boxes;
    const int64_t nbox = sweep_boxes(g, boxes).ntiles;                // the pattern is the structured stencil
    int64_t b = 4 * ((N + 1) + nnzb + N + N + nbf + nbox);         // rowptr, colidx, diag, level_rows, bc_type, boxes
    b += 8 * (N + 3 * nbf * m + 7 * n + nnzb * mm);                // K, bc/ghost values, 7 vectors, J
    if (m == 2 && desc->diff_method == DMX_DIFF_ANALYTIC) b += 8 * 3 * N; // saturation derivatives
    if (desc->material_id) b += 4 * (N + nbf);                       // material per cell / boundary face
    b += int64_t(sizeof(KrylovScalars)) + 8 * 8 + 4 + 8 * 4 * cap; // scalars, flag, partials
    if (batched)  // FP32 stencil operator and Dinv; FP64 Dinv, Krylov vectors, refinement residual
        b += 4 * (2 * g.dim * mm * N + N * mm) + 8 * (N * mm + 7 * n + n);
    if (batched == 2) b += 4 * mg_floats(g, int(m));                 // multigrid coarse levels
    *bytes = b;
    return DMX_OK;
}

int dmx_num_dofs(const dmx_context* ctx, int64_t* ncells, int32_t* m, int64_t* nnzb)
{
    if (!ctx) return DMX_ERR_INVALID;
    if (ncells) *ncells = ctx->N;
    if (m) *m = ctx->m;
    if (nnzb) *nnzb = ctx->nnzb;
    return DMX_OK;
}

int dmx_pattern(const dmx_context* ctx, int32_t* rowptr, int32_t* colidx)
{
    if (!ctx) return DMX_ERR_INVALID;
    if (rowptr) std::copy(ctx->h_rowptr.begin(), ctx->h_rowptr.end(), rowptr);
    if (colidx) std::copy(ctx->h_colidx.begin(), ctx->h_colidx.end(), colidx);
    return DMX_OK;
}

int dmx_level_schedule(const dmx_context* ctx, int32_t* nlevels, int32_t* level_ptr, int32_t* level_rows)
{
    if (!ctx) return DMX_ERR_INVALID;
    if (nlevels) *nlevels = ctx->nlev;
    if (level_ptr) std::copy(ctx->h_level_ptr.begin(), ctx->h_level_ptr.end(), level_ptr);
    if (level_rows) std::copy(ctx->h_level_rows.begin(), ctx->h_level_rows.end(), level_rows);
    return DMX_OK;
}

int dmx_set_solution(dmx_context* ctx, const double* u)
{
    if (!ctx || !u) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(hipMemcpyAsync(ctx->u, u, sizeof(double) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->haveState = true;
    ctx->jacValid = false;
    return DMX_OK;
}

int dmx_get_solution(const dmx_context* cctx, double* u)
{
    auto* ctx = const_cast<dmx_context*>(cctx);
    if (!ctx || !u) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(hipMemcpyAsync(u, ctx->u, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    return DMX_OK;
}

int dmx_get_volume_variables(dmx_context* ctx, double* p, double* rho_mob)
{
    if (!ctx || !p || !rho_mob || !ctx->haveState) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(launch_cell_state(ctx->stream, ctx->ph, ctx->u, ctx->vvp, ctx->vvu, nullptr, ctx->mat, ctx->N));
    DMX_HIP(hipMemcpyAsync(p, ctx->vvp, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    DMX_HIP(hipMemcpyAsync(rho_mob, ctx->vvu, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    return DMX_OK;
}

int dmx_set_prev_solution(dmx_context* ctx, const double* uprev)
{
    if (!ctx) return DMX_ERR_INVALID;
    if (!uprev) {
        ctx->havePrev = false;
        return DMX_OK;
    }
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(hipMemcpyAsync(ctx->uprev, uprev, sizeof(double) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->havePrev = true;
    return DMX_OK;
}

int dmx_reset_time_step(dmx_context* ctx)
{
    if (!ctx || !ctx->havePrev) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(launch_copy(ctx->stream, ctx->uprev, ctx->u, ctx->n));
    ctx->jacValid = false;
    return DMX_OK;
}

int dmx_advance_time_step(dmx_context* ctx)
{
    if (!ctx) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(launch_copy(ctx->stream, ctx->u, ctx->uprev, ctx->n));
    ctx->havePrev = true;
    return DMX_OK;
}

int dmx_set_time_step_size(dmx_context* ctx, double dt)
{
    if (!ctx || !(dt > 0.0)) return DMX_ERR_INVALID;
    ctx->dt = dt;
    return DMX_OK;
}

int dmx_assemble_jacobian_and_residual(dmx_context* ctx)
{
    if (!ctx || !ctx->haveState) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(launch_cell_state(ctx->stream, ctx->ph, ctx->u, ctx->vvp, ctx->vvu, ctx->vvd, ctx->mat, ctx->N));
    DMX_HIP(launch_assemble(ctx->stream, ctx->g, ctx->ph, asm_ptrs(ctx, true), true));
    ctx->jacValid = true;
    return DMX_OK;
}

int dmx_partial_reassembly_reset(dmx_context* ctx)
{
    if (!ctx) return DMX_ERR_INVALID;
    if (ctx->desc.diff_method == DMX_DIFF_ANALYTIC) {
        ctx->err = "partial reassembly needs the numeric Jacobian (its column reset presumes column-wise assembly)";
        return DMX_ERR_UNSUPPORTED;
    }
    DMX_HIP(hipSetDevice(ctx->device));
    if (!ctx->prDist) DMX_TRY(dalloc(ctx, &ctx->prDist, ctx->N));
    if (!ctx->prGreen) DMX_TRY(dalloc(ctx, &ctx->prGreen, ctx->N));
    DMX_HIP(hipMemsetAsync(ctx->prDist, 0, sizeof(double) * ctx->N, ctx->stream));
    DMX_HIP(hipMemsetAsync(ctx->prGreen, 0, ctx->N, ctx->stream));   // all red
    return DMX_OK;
}

int dmx_partial_reassembly_update(dmx_context* ctx, double threshold, int64_t* green_cells)
{
    if (!ctx || !ctx->prGreen) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(launch_partial_colors(ctx->stream, ctx->N, ctx->m, ctx->u_last, ctx->u, threshold, ctx->prDist, ctx->prGreen,
                                  ctx->scal + 2, ctx->rw));
    DMX_HIP(hipMemcpyAsync(ctx->scal_host + 2, ctx->scal + 2, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    if (green_cells) *green_cells = int64_t(ctx->scal_host[2]);
    return DMX_OK;
}

int dmx_get_partial_colors(const dmx_context* cctx, uint8_t* green)
{
    auto* ctx = const_cast<dmx_context*>(cctx);
    if (!ctx || !green || !ctx->prGreen) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(hipMemcpyAsync(green, ctx->prGreen, ctx->N, hipMemcpyDeviceToHost, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    return DMX_OK;
}

int dmx_assemble_jacobian_and_residual_partial(dmx_context* ctx)
{
    if (!ctx || !ctx->haveState) return DMX_ERR_INVALID;
    if (!ctx->prGreen) {
        ctx->err = "partial reassembly: dmx_partial_reassembly_reset first";
        return DMX_ERR_INVALID;
    }
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(launch_cell_state(ctx->stream, ctx->ph, ctx->u, ctx->vvp, ctx->vvu, ctx->vvd, ctx->mat, ctx->N));
    AsmPtrs P = asm_ptrs(ctx, true);
    P.green = ctx->prGreen;
    DMX_HIP(launch_assemble(ctx->stream, ctx->g, ctx->ph, P, true));
    ctx->jacValid = true;
    return DMX_OK;
}

int dmx_assemble_residual(dmx_context* ctx)
{
    if (!ctx || !ctx->haveState) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(launch_cell_state(ctx->stream, ctx->ph, ctx->u, ctx->vvp, ctx->vvu, nullptr, ctx->mat, ctx->N));
    DMX_HIP(launch_assemble(ctx->stream, ctx->g, ctx->ph, asm_ptrs(ctx, false), false));
    return DMX_OK;
}

int dmx_get_residual(const dmx_context* cctx, double* r)
{
    auto* ctx = const_cast<dmx_context*>(cctx);
    if (!ctx || !r) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(hipMemcpyAsync(r, ctx->r, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    return DMX_OK;
}

int dmx_get_jacobian(const dmx_context* cctx, double* values)
{
    auto* ctx = const_cast<dmx_context*>(cctx);
    if (!ctx || !values) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(hipMemcpyAsync(values, ctx->J, sizeof(double) * ctx->nnzb * ctx->m * ctx->m, hipMemcpyDeviceToHost, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    return DMX_OK;
}

int dmx_get_update(const dmx_context* cctx, double* du)
{
    auto* ctx = const_cast<dmx_context*>(cctx);
    if (!ctx || !du) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(hipMemcpyAsync(du, ctx->du, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    return DMX_OK;
}

int dmx_residual_norm(dmx_context* ctx, double* norm)
{
    if (!ctx || !norm) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    return device_norm(ctx, ctx->r, norm);
}

void dmx_default_linear_params(dmx_linear_params* p)
{
    if (!p) return;
    p->reduction = 1e-13;   // LinearSolver.ResidualReduction tree default (linearsolverparameters.hh:60)
    p->max_iter = 250;
    p->precond = DMX_PRECOND_ILU0;
    p->precision = DMX_PRECISION_FP64;
    p->inner_reduction = 1e-4;
    p->max_refinements = 20;
    p->ilu_sweep = 0;
}

int dmx_apply_preconditioner(dmx_context* ctx, int32_t precond, int32_t precision, const double* d, double* v)
{
    if (!ctx || !d || !v) return DMX_ERR_INVALID;
    if (!ctx->jacValid) {
        ctx->err = "preconditioner without an assembled Jacobian";
        return DMX_ERR_INVALID;
    }
    if (precision != DMX_PRECISION_FP64 && precision != DMX_PRECISION_MIXED) {
        ctx->err = "unknown precision";
        return DMX_ERR_INVALID;
    }
    DMX_HIP(hipSetDevice(ctx->device));
    const bool f32 = precision == DMX_PRECISION_MIXED;
    hipStream_t s = ctx->stream;
    DMX_TRY(ensure_krylov(ctx, f32));
    DMX_TRY(precond_setup(ctx, precond, f32));
    // the apply kernels return at once on a finished solve: clear the device scalars
    DMX_HIP(hipMemsetAsync(ctx->ks, 0, sizeof(KrylovScalars), s));
    if (!f32) {
        DMX_HIP(hipMemcpyAsync(ctx->kv64[2], d, sizeof(double) * ctx->n, hipMemcpyHostToDevice, s));
        DMX_TRY(precond_apply<double>(ctx, precond, ctx->LU, ctx->Dinv64, ctx->kv64[2], ctx->kv64[5]));
        DMX_HIP(hipMemcpyAsync(v, ctx->kv64[5], sizeof(double) * ctx->n, hipMemcpyDeviceToHost, s));
    } else {
        if (!ctx->ir_r) DMX_TRY(dalloc(ctx, &ctx->ir_r, ctx->n));
        DMX_HIP(hipMemcpyAsync(ctx->ir_r, d, sizeof(double) * ctx->n, hipMemcpyHostToDevice, s));
        DMX_HIP(launch_convert_f32(s, ctx->ir_r, ctx->kv32[2], ctx->n));
        DMX_TRY(precond_apply<float>(ctx, precond, ctx->LU32, ctx->Dinv32, ctx->kv32[2], ctx->kv32[5]));
        DMX_HIP(launch_fill(s, ctx->ir_r, 0.0, ctx->n));
        DMX_HIP(launch_add_to<float>(s, ctx->n, ctx->kv32[5], 1.0, ctx->ir_r));
        DMX_HIP(hipMemcpyAsync(v, ctx->ir_r, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, s));
    }
    DMX_HIP(hipStreamSynchronize(s));
    if (precond == DMX_PRECOND_ILU0) {
        int32_t failed = 0;
        DMX_HIP(hipMemcpy(&failed, ctx->fail, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (failed & 2) {
            ctx->err = "ILU(0) sweep: a row waited past the spin limit";
            return DMX_ERR_DEVICE;
        }
    }
    return DMX_OK;
}

int dmx_linear_solve(dmx_context* ctx, const dmx_linear_params* p, dmx_solve_info* info)
{
    if (!ctx || !p || !info) return DMX_ERR_INVALID;
    if (!ctx->jacValid) {
        ctx->err = "linear solve without an assembled Jacobian";
        return DMX_ERR_INVALID;
    }
    DMX_HIP(hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    *info = dmx_solve_info{};
    const bool mixed = p->precision == DMX_PRECISION_MIXED;
    ctx->iluSweep = p->ilu_sweep;
    DMX_TRY(precond_setup(ctx, p->precond, mixed));
    int conv = 0, its = 0;
    if (!mixed) {
        DMX_TRY(ensure_krylov(ctx, false));
        DMX_TRY(bicgstab<double>(ctx, ctx->J, ctx->LU, ctx->Dinv64, p->precond, ctx->r, 1.0, ctx->du, p->reduction,
                                 p->max_iter, &conv, &its));
        info->iterations = its;
        info->converged = conv;
        info->reduction = ctx->ks_host->def0 > 0 ? ctx->ks_host->norm / ctx->ks_host->def0 : 0.0;
    } else {
        // FP32 BiCGSTAB inside FP64 iterative refinement: du += A32^-1 (r - J du) until
        // ||r - J du|| <= reduction * ||r|| (all residuals in FP64)
        DMX_TRY(ensure_krylov(ctx, true));
        const int mm = ctx->m * ctx->m;
        if (!ctx->J32) DMX_TRY(dalloc(ctx, &ctx->J32, ctx->nnzb * mm));
        if (!ctx->ir_r) DMX_TRY(dalloc(ctx, &ctx->ir_r, ctx->n));
        DMX_HIP(launch_convert_f32(ctx->stream, ctx->J, ctx->J32, ctx->nnzb * mm));
        DMX_HIP(launch_fill(ctx->stream, ctx->du, 0.0, ctx->n));
        double normb = 0.0;
        DMX_TRY(device_norm(ctx, ctx->r, &normb));
        DMX_HIP(launch_copy(ctx->stream, ctx->r, ctx->ir_r, ctx->n));
        double res = normb;
        int k = 0;
        conv = normb == 0.0 ? 1 : 0;
        while (!conv && k < p->max_refinements) {
            int c32 = 0, i32 = 0;
            float* x32 = ctx->kv32[6];
            // the FP32 solve sees the residual normalized to unit 2-norm (no FP32 underflow of the
            // dot products near convergence); the correction is scaled back in FP64
            DMX_TRY(bicgstab<float>(ctx, ctx->J32, ctx->LU32, ctx->Dinv32, p->precond, ctx->ir_r, 1.0 / res, x32,
                                    p->inner_reduction, p->max_iter, &c32, &i32));
            its += i32;
            ++k;
            if (ctx->ks_host->status != 0 && i32 == 0) break;
            DMX_HIP(launch_add_to<float>(ctx->stream, ctx->n, x32, res, ctx->du));
            DMX_HIP(launch_residual(ctx->stream, ctx->m, ctx->N, ctx->rowptr, ctx->colidx, ctx->J, ctx->du, ctx->r, ctx->ir_r));
            DMX_TRY(device_norm(ctx, ctx->ir_r, &res));
            if (!(res == res)) break;
            if (res <= p->reduction * normb) conv = 1;
        }
        info->iterations = its;
        info->refinements = k;
        info->converged = conv;
        info->reduction = normb > 0 ? res / normb : 0.0;
    }
    info->elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!info->converged) {
        ctx->err = "linear solver did not converge";
        return DMX_NUMERICAL_PROBLEM;
    }
    return DMX_OK;
}

int dmx_newton_update(dmx_context* ctx, double* shift)
{
    if (!ctx) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    // keep the old iterate: swap buffers, then u = u_last - du
    std::swap(ctx->u, ctx->u_last);
    DMX_HIP(launch_newton_update(ctx->stream, ctx->n, ctx->u_last, ctx->du, ctx->u, ctx->scal, ctx->rw));
    DMX_HIP(hipMemcpyAsync(ctx->scal_host, ctx->scal, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    if (shift) *shift = ctx->scal_host[0];
    ctx->jacValid = false;
    return DMX_OK;
}

void dmx_default_newton_params(dmx_newton_params* p)
{
    if (!p) return;
    Dumux::Gpu::NewtonParameters d;
    p->max_rel_shift = d.maxRelativeShift;
    p->min_steps = d.minSteps;
    p->max_steps = d.maxSteps;
    p->max_dt_divisions = d.maxTimeStepDivisions;
    p->retry_factor = d.retryTimeStepReductionFactor;
    p->linear_pde = 0;
    p->verbosity = 0;
    p->adaptive_dt = 0;
    p->target_steps = d.targetSteps;
    p->max_dt = 0.0;
    p->partial_reassembly = 0;
    p->reassembly_min_threshold = 0.0;   // 1e-1 * max_rel_shift (newtonsolver.hh:1242)
    p->reassembly_max_threshold = 0.0;   // 1e2 * max_rel_shift (:1243)
    p->reassembly_shift_weight = 1e-3;   // (:1244)
    dmx_default_linear_params(&p->linear);
    p->linear.reduction = 1e-6;  // NewtonSolver sets LinearSolver.ResidualReduction (newtonsolver.hh:232)
}

} // extern "C"

namespace {

Dumux::Gpu::NewtonParameters newton_params(const dmx_newton_params* p)
{
    Dumux::Gpu::NewtonParameters np;
    np.maxRelativeShift = p->max_rel_shift;
    np.minSteps = p->min_steps;
    np.maxSteps = p->max_steps;
    np.maxTimeStepDivisions = p->max_dt_divisions;
    np.retryTimeStepReductionFactor = p->retry_factor;
    np.targetSteps = p->target_steps;
    np.verbosity = p->verbosity;
    np.enablePartialReassembly = p->partial_reassembly != 0;
    np.reassemblyMinThreshold = p->reassembly_min_threshold;
    np.reassemblyMaxThreshold = p->reassembly_max_threshold;
    np.reassemblyShiftWeight = p->reassembly_shift_weight;
    return np;
}

template <class Newton>
void fill_info(const Newton& nw, const Dumux::Gpu::LinearSolver& ls, dmx_newton_info* info)
{
    info->newton_steps = nw.totalSucceededIterations();
    info->wasted_steps = nw.totalWastedIterations();
    info->linear_iterations = ls.iterations();
    info->refinements = ls.refinements();
    info->dt_retries = nw.dtRetries();
    info->last_shift = nw.lastShift();
    info->t_assemble = nw.assembleTime();
    info->t_solve = nw.solveTime();
    info->t_update = nw.updateTime();
    info->reassembled_cells = nw.reassembledCells();
}

} // namespace

extern "C" {

int dmx_newton_solve(dmx_context* ctx, const dmx_newton_params* p, dmx_newton_info* info)
{
    if (!ctx || !p || !info) return DMX_ERR_INVALID;
    *info = dmx_newton_info{};
    if (p->partial_reassembly && ctx->desc.diff_method == DMX_DIFF_ANALYTIC) {
        ctx->err = "partial reassembly needs the numeric Jacobian";
        return DMX_ERR_UNSUPPORTED;
    }
    try {
        auto assembler = std::make_shared<Dumux::Gpu::FVAssembler>(ctx);
        assembler->setStationary(!ctx->havePrev);
        auto ls = std::make_shared<Dumux::Gpu::LinearSolver>(p->linear);
        if (p->linear_pde) {
            // LinearPDESolver::apply (linear/pdesolver.hh:107-170)
            Dumux::Gpu::check(dmx_assemble_jacobian_and_residual(ctx), ctx, "assemble");
            if (!ls->solve(assembler->jacobian(), assembler->update(), assembler->residual())) {
                info->converged = 0;
                info->linear_iterations = ls->iterations();
                return DMX_NUMERICAL_PROBLEM;
            }
            double shift = 0.0;
            Dumux::Gpu::check(dmx_newton_update(ctx, &shift), ctx, "update");
            info->converged = 1;
            info->newton_steps = 1;
            info->last_shift = shift;
            info->linear_iterations = ls->iterations();
            info->refinements = ls->refinements();
            return DMX_OK;
        }
        Dumux::Gpu::DeviceNewtonSolver nw(assembler, ls, newton_params(p));
        const bool ok = nw.apply(assembler->solution());
        fill_info(nw, *ls, info);
        info->converged = ok ? 1 : 0;
        return ok ? DMX_OK : DMX_NUMERICAL_PROBLEM;
    } catch (const Dumux::Gpu::NumericalProblem& e) {
        ctx->err = e.what();
        return DMX_NUMERICAL_PROBLEM;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return DMX_ERR_DEVICE;
    }
}

int dmx_time_loop(dmx_context* ctx, int32_t nsteps, double dt, const dmx_newton_params* p, dmx_newton_info* info)
{
    if (!ctx || !p || !info || nsteps < 0 || !(dt > 0.0)) return DMX_ERR_INVALID;
    *info = dmx_newton_info{};
    if (p->partial_reassembly && ctx->desc.diff_method == DMX_DIFF_ANALYTIC) {
        ctx->err = "partial reassembly needs the numeric Jacobian";
        return DMX_ERR_UNSUPPORTED;
    }
    try {
        // the loop starts from the current solution as the old time level (xOld = x,
        // test/porousmediumflow/2p/incompressible/main.cc:102), whatever an earlier loop left
        Dumux::Gpu::check(dmx_advance_time_step(ctx), ctx, "init prev");
        auto assembler = std::make_shared<Dumux::Gpu::FVAssembler>(ctx);
        auto ls = std::make_shared<Dumux::Gpu::LinearSolver>(p->linear);
        Dumux::Gpu::DeviceNewtonSolver nw(assembler, ls, newton_params(p));
        Dumux::Gpu::TimeLoopState tl;
        tl.tEnd = nsteps * dt;
        const double maxDt = p->max_dt > 0.0 ? p->max_dt : tl.tEnd;
        double dtNext = std::min(dt, maxDt);
        while (tl.time < tl.tEnd) {
            tl.dt = p->adaptive_dt ? std::min(dtNext, std::min(maxDt, tl.tEnd - tl.time)) : std::min(dt, tl.tEnd - tl.time);
            assembler->setTimeStepSize(tl.dt);
            nw.solve(assembler->solution(), tl);
            if (p->adaptive_dt) dtNext = nw.suggestTimeStepSize(tl.dt);
            tl.time += tl.dt;
            assembler->advanceTimeStep();
            info->time_steps++;
        }
        fill_info(nw, *ls, info);
        info->converged = 1;
        DMX_HIP(hipStreamSynchronize(ctx->stream));
        return DMX_OK;
    } catch (const Dumux::Gpu::NumericalProblem& e) {
        ctx->err = e.what();
        info->converged = 0;
        return DMX_NUMERICAL_PROBLEM;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return DMX_ERR_DEVICE;
    }
}

int dmx_time_kernel(dmx_context* ctx, int32_t kernel, int32_t reps, double* avg_ms, double* bytes)
{
    if (!ctx || reps < 1 || !avg_ms || !bytes) return DMX_ERR_INVALID;
    if (!ctx->jacValid && kernel != DMX_KERNEL_CELL_STATE && kernel != DMX_KERNEL_RESIDUAL && kernel != DMX_KERNEL_ASSEMBLE) {
        ctx->err = "time_kernel needs an assembled Jacobian";
        return DMX_ERR_INVALID;
    }
    DMX_HIP(hipSetDevice(ctx->device));
    const double N = double(ctx->N), m = ctx->m, nnzb = double(ctx->nnzb), n = double(ctx->n);
    const double inst = ctx->havePrev ? 1.0 : 0.0;
    const bool fused = kernel >= DMX_KERNEL_BJ_SPMV_DOT_FP32 && kernel <= DMX_KERNEL_BJ_ITERATION_FP32;
    const bool need32 = kernel == DMX_KERNEL_SPMV_FP32 || kernel == DMX_KERNEL_JACOBI_APPLY_FP32 ||
                        kernel == DMX_KERNEL_BICG_HALF_FP32 || fused;
    const bool need64 = kernel == DMX_KERNEL_SPMV_FP64 || kernel == DMX_KERNEL_ILU0_APPLY_FP64;
    if ((need32 && (!ctx->kv32[0] || !ctx->J32)) || (need64 && !ctx->kv64[0]) ||
        (kernel == DMX_KERNEL_ILU0_APPLY_FP64 && !ctx->LU) ||
        ((kernel == DMX_KERNEL_JACOBI_APPLY_FP32 || fused) && !ctx->Dinv32) ||
        (kernel == DMX_KERNEL_REFINE_RESIDUAL && !ctx->ir_r)) {
        ctx->err = "time_kernel: run a linear solve of this precision/preconditioner first";
        return DMX_ERR_INVALID;
    }
    hipStream_t s = ctx->stream;
    {
        // benign scalars: never converged (target 0), no breakdown, not the first iteration
        KrylovScalars k0{};
        k0.rho = k0.rho_new = 1.0;
        k0.alpha = k0.omega = 1e-3;
        k0.norm = k0.def0 = 1.0;
        k0.reduction_target = 0.0;
        k0.half_steps = 2;
        *ctx->ks_host = k0;
        DMX_HIP(hipMemcpyAsync(ctx->ks, ctx->ks_host, sizeof(KrylovScalars), hipMemcpyHostToDevice, s));
        DMX_HIP(hipStreamSynchronize(s));
    }
    // the ILU(0) apply is timed in the form the last solve used
    const bool sweep = kernel == DMX_KERNEL_ILU0_APPLY_FP64 && ctx->sweepFits && ctx->iluSweep == 0 && ctx->sweepTmp64;
    hipEvent_t e0, e1;
    DMX_HIP(hipEventCreate(&e0));
    DMX_HIP(hipEventCreate(&e1));
    auto once = [&]() -> hipError_t {
        switch (kernel) {
        case DMX_KERNEL_ASSEMBLE: return launch_assemble(s, ctx->g, ctx->ph, asm_ptrs(ctx, true), true);
        case DMX_KERNEL_CELL_STATE:
            return launch_cell_state(s, ctx->ph, ctx->u, ctx->vvp, ctx->vvu, ctx->vvd, ctx->mat, ctx->N);
        case DMX_KERNEL_RESIDUAL: return launch_assemble(s, ctx->g, ctx->ph, asm_ptrs(ctx, false), false);
        case DMX_KERNEL_SPMV_FP64:
            return launch_spmv_tile<double>(s, ctx->m, ctx->N, ctx->rowptr, ctx->colidx, ctx->J, ctx->kv64[5], ctx->kv64[3]);
        case DMX_KERNEL_SPMV_FP32:
            return launch_spmv_tile<float>(s, ctx->m, ctx->N, ctx->rowptr, ctx->colidx, ctx->J32, ctx->kv32[5], ctx->kv32[3]);
        case DMX_KERNEL_BJ_SPMV_DOT_FP32:
        case DMX_KERNEL_BJ_SPMV_DOTS_FP32:
        case DMX_KERNEL_BJ_ITERATION_FP32: {
            float* vec[7] = {ctx->kv32[0], ctx->kv32[1], ctx->kv32[2], ctx->kv32[3], ctx->kv32[4], ctx->kv32[5], ctx->kv32[6]};
            if (kernel == DMX_KERNEL_BJ_ITERATION_FP32)
                return bj_iteration<float>(s, ctx->m, ctx->N, ctx->rowptr, ctx->colidx, ctx->J32, ctx->Dinv32, vec, ctx->ks, ctx->rw, 1 << 28);
            return bj_kernel<float>(s, kernel == DMX_KERNEL_BJ_SPMV_DOT_FP32 ? 1 : 3, ctx->m, ctx->N, ctx->rowptr, ctx->colidx,
                                    ctx->J32, ctx->Dinv32, vec, ctx->ks, ctx->rw, 1 << 28);
        }
        case DMX_KERNEL_ILU0_APPLY_FP64:
            if (sweep)
                return launch_ilu0_sweeps<double>(s, ctx->m, ctx->g, ctx->tg, ctx->boxOrder, ctx->rowptr, ctx->colidx,
                                                  ctx->diag, ctx->LU, ctx->sweepGrid, ctx->kv64[2], ctx->sweepTmp64,
                                                  ctx->kv64[5], ctx->sweepTickets, ctx->fail, ctx->ks);
            for (int l = 0; l < ctx->nlev; ++l) {
                const int32_t beg = ctx->h_level_ptr[l], cnt = ctx->h_level_ptr[l + 1] - beg;
                hipError_t e = launch_ilu0_lower_level<double>(s, ctx->m, ctx->rowptr, ctx->colidx, ctx->LU,
                                                               ctx->level_rows + beg, cnt, ctx->kv64[2], ctx->kv64[5], nullptr);
                if (e != hipSuccess) return e;
            }
            for (int l = ctx->nlev - 1; l >= 0; --l) {
                const int32_t beg = ctx->h_level_ptr[l], cnt = ctx->h_level_ptr[l + 1] - beg;
                hipError_t e = launch_ilu0_upper_level<double>(s, ctx->m, ctx->rowptr, ctx->colidx, ctx->diag, ctx->LU,
                                                               ctx->level_rows + beg, cnt, ctx->kv64[5], nullptr);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        case DMX_KERNEL_JACOBI_APPLY_FP32:
            return launch_jacobi_apply<float>(s, ctx->m, ctx->N, ctx->Dinv32, ctx->kv32[2], ctx->kv32[5], nullptr);
        case DMX_KERNEL_BICG_HALF_FP32:
            return bicg_half1<float>(s, ctx->n, ctx->kv32[5], ctx->kv32[3], ctx->kv32[6], ctx->kv32[4], ctx->ks, ctx->rw);
        case DMX_KERNEL_NEWTON_UPDATE:
            return launch_newton_update(s, ctx->n, ctx->u, ctx->du, ctx->vvp, ctx->scal + 2, ctx->rw);
        case DMX_KERNEL_REFINE_RESIDUAL:
            return launch_residual(s, ctx->m, ctx->N, ctx->rowptr, ctx->colidx, ctx->J, ctx->du, ctx->r, ctx->ir_r);
        default: return hipErrorInvalidValue;
        }
    };
    hipError_t e = once();  // warm-up
    if (e != hipSuccess) {
        ctx->err = std::string("time_kernel: ") + hipGetErrorString(e);
        return DMX_ERR_DEVICE;
    }
    DMX_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) {
        e = once();
        if (e != hipSuccess) break;
    }
    DMX_HIP(hipEventRecord(e1, s));
    DMX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DMX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) {
        ctx->err = std::string("time_kernel: ") + hipGetErrorString(e);
        return DMX_ERR_DEVICE;
    }
    *avg_ms = double(ms) / reps;
    if (sweep) {
        int32_t failed = 0;
        DMX_HIP(hipMemcpy(&failed, ctx->fail, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (failed & 2) {
            ctx->err = "time_kernel: ILU(0) sweep row waited past the spin limit";
            return DMX_ERR_DEVICE;
        }
    }
    if (fused) {
        DMX_HIP(hipMemcpyAsync(ctx->ks_host, ctx->ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, s));
        DMX_HIP(hipStreamSynchronize(s));
        if (ctx->ks_host->done) {
            ctx->err = "time_kernel: the fused kernels stopped early (breakdown on the timing data)";
            return DMX_NUMERICAL_PROBLEM;
        }
    }
    const double mm = m * m;
    switch (kernel) {
    case DMX_KERNEL_ASSEMBLE: *bytes = 8 * N * (m + m * inst + 1) + 8 * n + 8 * mm * nnzb; break;
    case DMX_KERNEL_CELL_STATE: *bytes = 8 * n + 2 * 8 * n; break;
    case DMX_KERNEL_RESIDUAL: *bytes = 8 * N * (m + m * inst + 1) + 8 * n; break;
    case DMX_KERNEL_SPMV_FP64: *bytes = nnzb * (8 * mm + 4) + 4 * (N + 1) + 2 * 8 * n; break;
    case DMX_KERNEL_SPMV_FP32: *bytes = nnzb * (4 * mm + 4) + 4 * (N + 1) + 2 * 4 * n; break;
    case DMX_KERNEL_ILU0_APPLY_FP64: *bytes = nnzb * (8 * mm + 4) + 2 * 8 * n; break;
    case DMX_KERNEL_JACOBI_APPLY_FP32: *bytes = 4 * mm * N + 2 * 4 * n; break;
    case DMX_KERNEL_BICG_HALF_FP32: *bytes = 6 * 4 * n; break;
    case DMX_KERNEL_NEWTON_UPDATE: *bytes = 3 * 8 * n; break;
    case DMX_KERNEL_REFINE_RESIDUAL: *bytes = nnzb * (8 * mm + 4) + 4 * (N + 1) + 3 * 8 * n; break;
    // SpMV (matrix + indices + gathered y + written result) plus the vectors of the fused dots
    case DMX_KERNEL_BJ_SPMV_DOT_FP32: *bytes = nnzb * (4 * mm + 4) + 4 * (N + 1) + 3 * 4 * n; break;
    case DMX_KERNEL_BJ_SPMV_DOTS_FP32: *bytes = nnzb * (4 * mm + 4) + 4 * (N + 1) + 3 * 4 * n; break;
    // A: r,v,p,Dinv -> p,y; B: SpMV+rt; C: x,r,y,v,Dinv -> x,r,y; D: SpMV+r; E: x,r,y,t,rt -> x,r
    case DMX_KERNEL_BJ_ITERATION_FP32:
        *bytes = (5 * 4 * n + 4 * mm * N) + (nnzb * (4 * mm + 4) + 4 * (N + 1) + 3 * 4 * n) +
                 (7 * 4 * n + 4 * mm * N) + (nnzb * (4 * mm + 4) + 4 * (N + 1) + 3 * 4 * n) + 7 * 4 * n;
        break;
    default: *bytes = 0;
    }
    return DMX_OK;
}

int dmx_synchronize(dmx_context* ctx)
{
    if (!ctx) return DMX_ERR_INVALID;
    DMX_HIP(hipSetDevice(ctx->device));
    DMX_HIP(hipStreamSynchronize(ctx->stream));
    return DMX_OK;
}

} // extern "C"
