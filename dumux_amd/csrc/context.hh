// Per-problem device context of the C ABI (one problem, one HIP stream), shared by the single
// problem entry points (capi.hip) and the batched executor (batch.hip).
#pragma once
#include "setup.hh"
#include "kernels.hh"

#include <string>
#include <vector>

using namespace dmx;

struct dmx_context {
    int device = 0;
    hipStream_t stream = nullptr;
    dmx_problem_desc desc{};
    GridDesc g{};
    Phys ph{};
    int m = 1;
    int64_t N = 0, nbf = 0, nnzb = 0, n = 0;  // n = N*m
    std::vector<int32_t> h_rowptr, h_colidx, h_level_ptr, h_level_rows, h_diag;
    int nlev = 0;

    int32_t *rowptr = nullptr, *colidx = nullptr, *diag = nullptr, *level_rows = nullptr, *bct = nullptr;
    double *K = nullptr, *bcv = nullptr, *ghost_p = nullptr, *ghost_up = nullptr;
    double *u = nullptr, *u_last = nullptr, *uprev = nullptr, *vvp = nullptr, *vvu = nullptr;
    double* vvd = nullptr;              // 2p analytic Jacobian: 3 saturation derivatives per cell
    int32_t *mat = nullptr, *bmat = nullptr;  // element-wise materials: per cell, per boundary face
    double *J = nullptr, *r = nullptr, *du = nullptr;
    double *LU = nullptr, *Dinv64 = nullptr;
    float *J32 = nullptr, *Dinv32 = nullptr, *LU32 = nullptr;
    float* S32 = nullptr;               // batched executor: scaled FP32 operator, stencil layout
    double* kvb = nullptr;              // batched executor: Krylov vectors (7 FP64 or up to 14 FP32)
    float* mgBuf = nullptr;             // batched executor: multigrid coarse levels
    std::vector<MgLevel> mgLev;         // single-problem multigrid levels (DMX_PRECOND_AMG)
    float* mg1Buf = nullptr;            // their operators and vectors
    MgLevel* mg1LevDev = nullptr;       // mgLev in device memory
    int mg1Tail = 0;                    // first level of the one-workgroup tail
    double* prDist = nullptr;           // partial reassembly: distance from the last linearization per cell
    uint8_t* prGreen = nullptr;         // partial reassembly: 1 = green (keep the Jacobian column)
    double* kv64[7] = {};
    float* kv32[7] = {};
    double *ir_r = nullptr, *ir_x = nullptr;
    KrylovScalars* ks = nullptr;
    KrylovScalars* ks_host = nullptr;   // pinned
    double* scal = nullptr;             // device scalars: [0] shift, [1] norm
    double* scal_host = nullptr;        // pinned
    int32_t* fail = nullptr;
    // ILU(0) sweeps in one launch each: boxes of cells drawn in box-level order from two
    // counters, the forward sweep's result in a pattern-filled buffer of the solve's precision
    bool sweepFits = false;             // the pattern is the structured stencil of g
    int iluSweep = 0;                   // 0 one launch per sweep, 1 one launch per level
    int sweepGrid = 256;
    TileGeom tg{};
    std::vector<int32_t> h_boxOrder;
    int32_t* boxOrder = nullptr;
    uint32_t* sweepTickets = nullptr;
    double* sweepTmp64 = nullptr;
    float* sweepTmp32 = nullptr;
    ReduceWork rw{};

    // captured graphs of kGraphIters fused block-Jacobi BiCGSTAB iterations ([0] FP32, [1] FP64)
    hipGraph_t bjGraph[2] = {nullptr, nullptr};
    hipGraphExec_t bjExec[2] = {nullptr, nullptr};
    int bjMaxit[2] = {-1, -1};

    int64_t devBytes = 0;    // device memory held by the context
    double dt = 1.0;
    bool haveState = false;
    bool havePrev = false;
    bool jacValid = false;   // J matches u
    std::string err;
};

namespace dmx {

#define DMX_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return DMX_ERR_DEVICE;                                                      \
        }                                                                               \
    } while (0)

template <typename T>
inline int dalloc(dmx_context* ctx, T** p, int64_t count)
{
    if (count <= 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * size_t(count));
    if (e != hipSuccess) {
        ctx->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        *p = nullptr;
        return DMX_ERR_NOMEM;
    }
    ctx->devBytes += int64_t(sizeof(T)) * count;
    return DMX_OK;
}

#define DMX_TRY(expr)               \
    do {                            \
        int rc_ = (expr);           \
        if (rc_ != DMX_OK) return rc_; \
    } while (0)


} // namespace dmx
