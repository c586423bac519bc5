// Batched executor: many independent problems (contexts) advanced together, one set of kernel
// launches per Newton phase for all of them.
//
// Layout: every problem keeps its device arrays in its own context; the batch holds a device
// table of per-problem descriptors (BatchProblem) and, per kind (dim, unknowns per cell), a tile
// table: tile = 256 consecutive cells / block rows of one problem, tiles never straddle
// problems. Every batched kernel runs one workgroup per tile and looks its problem up in the
// table, so the block-diagonal system of all problems is concatenated logically without
// copying, and problems can join or leave between phases (continuous batching).
//
// Per problem: its own Newton state (host), its own time level, its own BiCGSTAB scalars and
// convergence flags (device). A reduction over a problem sums that problem's tile partials in
// tile order, and tiles start at the problem's first row, so a problem's results are bitwise
// the same alone and in any mix.
//
// Linear solves: FP32 BiCGSTAB with block Jacobi (the FP32 matrix and the inverted diagonal
// blocks come out of the assembly kernel) inside FP64 iterative refinement, as the
// single-problem DMX_PRECISION_MIXED path.
#include "context.hh"
#include "krylov_ops.hh"
#include "../../include/dumux_amd/newtonsolver.hh"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace dmx {

namespace {

constexpr int kBT = 256;           // cells per tile = threads per workgroup
constexpr int kMaxRowBlocks = 7;

enum : int32_t { F_NEWTON = 1, F_LIN = 2, F_ADVANCE = 4, F_RESET = 8 };

struct BatchProblem {
    GridDesc g;
    Phys ph;
    AsmPtrs P;                 // assembly pointers; P.J32 and P.Dinv32 are set
    double* vvp;
    double* vvu;
    const int32_t* rowptr;
    const int32_t* colidx;
    const float* J32;
    const float* Dinv32;
    const double* J;
    float* v[7];               // r, rt, p, v, t, y, x (FP32 Krylov vectors)
    double* r;                 // Newton residual
    double* du;                // Newton update
    double* ir;                // refinement residual
    double* u;
    double* u_last;
    double* uprev;
    KrylovScalars* ks;
    int64_t N, n;
    int32_t tile0, ntiles;     // this problem's partial slots: tile0 .. tile0 + ntiles - 1
};

struct BatchDyn {
    int32_t flags;
    int32_t pad;
    double dt;
    double scale;              // right-hand side scaling of the FP32 solve (1/||r||)
    double corr;               // scaling of the FP32 correction (||r||)
};

struct TileRef {
    int32_t prob;              // global problem index
    int32_t local;             // tile index inside the problem
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ T wg_sum(T v, T* lds4)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    const T r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double wg_max(double v, double* lds4)
{
    for (int o = 32; o > 0; o >>= 1) v = __builtin_fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = __builtin_fmax(__builtin_fmax(lds4[0], lds4[1]), __builtin_fmax(lds4[2], lds4[3]));
    __syncthreads();
    return r;
}

template <int NV>
__device__ __forceinline__ void put_partials(const double* part, ReduceWork w)
{
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) w.partial[size_t(k) * w.cap + blockIdx.x] = part[k];
}

// ---- assembly -------------------------------------------------------------------------------
template <int M>
__global__ void __launch_bounds__(kBT) k_b_state(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                                 const TileRef* __restrict__ tiles)
{
    const TileRef t = tiles[blockIdx.x];
    if (!(dyn[t.prob].flags & F_NEWTON)) return;
    const BatchProblem& p = probs[t.prob];
    const int64_t I = int64_t(t.local) * kBT + threadIdx.x;
    if (I >= p.N) return;
    cell_state<M>(p.ph, p.P.u, p.vvp, p.vvu, I);
}

template <int DIM, int M>
__global__ void __launch_bounds__(kBT) k_b_assemble(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                                    const TileRef* __restrict__ tiles)
{
    const TileRef t = tiles[blockIdx.x];
    const BatchDyn d = dyn[t.prob];
    if (!(d.flags & F_NEWTON)) return;
    const BatchProblem& p = probs[t.prob];
    const int64_t I = int64_t(t.local) * kBT + threadIdx.x;
    if (I >= p.N) return;
    AsmPtrs P = p.P;
    P.dt = d.dt;
    assemble_cell<DIM, M, true>(p.g, p.ph, P, I);
}

// ---- vector helpers -------------------------------------------------------------------------
// op 0: du = 0, ir = r (start of a linear solve)   1: uprev = u (advance)   2: u = uprev (reset)
template <int M>
__global__ void __launch_bounds__(kBT) k_b_vec(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                               const TileRef* __restrict__ tiles, int op, int32_t flag)
{
    const TileRef t = tiles[blockIdx.x];
    if (!(dyn[t.prob].flags & flag)) return;
    const BatchProblem& p = probs[t.prob];
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    if (i >= p.N) return;
    for (int a = 0; a < M; ++a) {
        const int64_t k = i * M + a;
        if (op == 0) { p.du[k] = 0.0; p.ir[k] = p.r[k]; }
        else if (op == 1) p.uprev[k] = p.u[k];
        else p.u[k] = p.uprev[k];
    }
}

// ||r||^2 (op 0) or ||ir||^2 (op 1) partials
template <int M>
__global__ void __launch_bounds__(kBT) k_b_norm2(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                                 const TileRef* __restrict__ tiles, int op, int32_t flag, ReduceWork w)
{
    __shared__ double lds4[4];
    const TileRef t = tiles[blockIdx.x];
    if (!(dyn[t.prob].flags & flag)) return;
    const BatchProblem& p = probs[t.prob];
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    double acc = 0.0;
    if (i < p.N) {
        const double* x = op == 0 ? p.r : p.ir;
        for (int a = 0; a < M; ++a) acc += x[i * M + a] * x[i * M + a];
    }
    const double part = wg_sum<double>(acc, lds4);
    put_partials<1>(&part, w);
}

// u_last = u; u = u - du; max relative shift partials (newtonsolver.hh:110-116, 543-604)
template <int M>
__global__ void __launch_bounds__(kBT) k_b_update(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                                  const TileRef* __restrict__ tiles, ReduceWork w)
{
    __shared__ double lds4[4];
    const TileRef t = tiles[blockIdx.x];
    if (!(dyn[t.prob].flags & F_NEWTON)) return;
    const BatchProblem& p = probs[t.prob];
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    double mx = 0.0;
    if (i < p.N) {
        for (int a = 0; a < M; ++a) {
            const int64_t k = i * M + a;
            const double u1 = p.u[k];
            double u2 = u1;
            u2 += -1.0 * p.du[k];
            p.u_last[k] = u1;
            p.u[k] = u2;
            const double d = __builtin_fabs(u1 - u2) / __builtin_fmax(1.0, __builtin_fabs(u1 + u2) * 0.5);
            mx = __builtin_fmax(mx, d);
        }
    }
    const double part = wg_max(mx, lds4);
    put_partials<1>(&part, w);
}

// du += corr * x32 (refinement correction)
template <int M>
__global__ void __launch_bounds__(kBT) k_b_correct(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                                   const TileRef* __restrict__ tiles)
{
    const TileRef t = tiles[blockIdx.x];
    const BatchDyn d = dyn[t.prob];
    if (!(d.flags & F_LIN)) return;
    const BatchProblem& p = probs[t.prob];
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    if (i >= p.N) return;
    for (int a = 0; a < M; ++a) p.du[i * M + a] += d.corr * double(p.v[6][i * M + a]);
}

// tile SpMV (one tile = 256 consecutive rows of one problem)
template <int M, typename T>
__device__ __forceinline__ bool tile_spmv(const BatchProblem& p, int64_t tileLocal, const T* __restrict__ A,
                                          const T* __restrict__ x, T* lds, T (&acc)[M])
{
    const int64_t r0 = tileLocal * kBT;
    const int64_t r1 = r0 + kBT < p.N ? r0 + kBT : p.N;
    const int32_t k0 = p.rowptr[r0], k1 = p.rowptr[r1];
    for (int32_t k = k0 + threadIdx.x; k < k1; k += kBT) {
        const int64_t j = p.colidx[k];
        const T* B = A + int64_t(k) * M * M;
        T Bv[M * M];
        for (int q = 0; q < M * M; ++q) Bv[q] = B[q];
        T xv[M];
        for (int b = 0; b < M; ++b) xv[b] = x[j * M + b];
        for (int a = 0; a < M; ++a) {
            T s = T(0);
            for (int b = 0; b < M; ++b) s += Bv[a * M + b] * xv[b];
            lds[(k - k0) * M + a] = s;
        }
    }
    __syncthreads();
    const int64_t row = r0 + threadIdx.x;
    for (int a = 0; a < M; ++a) acc[a] = T(0);
    if (row >= r1) return false;
    const int32_t e = p.rowptr[row + 1];
    for (int32_t k = p.rowptr[row]; k < e; ++k)
        for (int a = 0; a < M; ++a) acc[a] += lds[(k - k0) * M + a];
    return true;
}

// ir = r - J du (FP64) and ||ir||^2 partials
template <int M>
__global__ void __launch_bounds__(kBT) k_b_residual(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                                    const TileRef* __restrict__ tiles, ReduceWork w)
{
    __shared__ double lds[kBT * kMaxRowBlocks * M];
    __shared__ double lds4[4];
    const TileRef t = tiles[blockIdx.x];
    if (!(dyn[t.prob].flags & F_LIN)) return;
    const BatchProblem& p = probs[t.prob];
    double acc[M];
    double nr = 0.0;
    if (tile_spmv<M, double>(p, t.local, p.J, p.du, lds, acc)) {
        const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
        for (int a = 0; a < M; ++a) {
            const double v = p.r[i * M + a] - acc[a];
            p.ir[i * M + a] = v;
            nr += v * v;
        }
    }
    const double part = wg_sum<double>(nr, lds4);
    put_partials<1>(&part, w);
}

// ---- FP32 BiCGSTAB with block Jacobi, per problem scalars ------------------------------------
__device__ __forceinline__ bool lin_active(const BatchDyn& d, const KrylovScalars* ks)
{
    return (d.flags & F_LIN) && !ks->done;
}

template <int M>
__device__ __forceinline__ void jac(const float* __restrict__ Dinv, int64_t i, const float (&p)[M], float (&y)[M])
{
    for (int a = 0; a < M; ++a) {
        float s = 0.f;
        for (int b = 0; b < M; ++b) s += Dinv[i * M * M + a * M + b] * p[b];
        y[a] = s;
    }
}

template <int M>
__global__ void __launch_bounds__(kBT) k_b_init(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                                const TileRef* __restrict__ tiles, ReduceWork w)
{
    __shared__ float lds4[4];
    const TileRef t = tiles[blockIdx.x];
    const BatchDyn d = dyn[t.prob];
    if (!(d.flags & F_LIN)) return;
    const BatchProblem& p = probs[t.prob];
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    float acc = 0.f;
    if (i < p.N) {
        for (int a = 0; a < M; ++a) {
            const int64_t k = i * M + a;
            const float ri = float(p.ir[k] * d.scale);
            p.v[0][k] = ri;
            p.v[1][k] = ri;
            p.v[2][k] = 0.f;
            p.v[3][k] = 0.f;
            p.v[6][k] = 0.f;
            acc += ri * ri;
        }
    }
    const double part = double(wg_sum<float>(acc, lds4));
    put_partials<1>(&part, w);
}

template <int M>
__global__ void __launch_bounds__(kBT) k_b_a(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                             const TileRef* __restrict__ tiles)
{
    const TileRef t = tiles[blockIdx.x];
    const BatchProblem& p = probs[t.prob];
    if (!lin_active(dyn[t.prob], p.ks)) return;
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    if (i >= p.N) return;
    const KrylovScalars* ks = p.ks;
    const bool first = ks->half_steps == 0;
    const double beta = (ks->norm == 0.0) ? 0.0 : (ks->rho_new / ks->rho) * (ks->alpha / ks->omega);
    const float om = float(ks->omega), bt = float(beta);
    float pv[M], yv[M];
    for (int a = 0; a < M; ++a) {
        const int64_t k = i * M + a;
        const float ri = p.v[0][k];
        if (first) {
            pv[a] = ri;
        } else {
            float q = p.v[2][k];
            q += -om * p.v[3][k];
            q *= bt;
            q += ri;
            pv[a] = q;
        }
        p.v[2][k] = pv[a];
    }
    jac<M>(p.Dinv32, i, pv, yv);
    for (int a = 0; a < M; ++a) p.v[5][i * M + a] = yv[a];
}

// B (which = 0): v = A y, <rt, v>;  D (which = 1): t = A y, <t, r>, <t, t>
template <int M>
__global__ void __launch_bounds__(kBT) k_b_spmv(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                                const TileRef* __restrict__ tiles, ReduceWork w, int which)
{
    __shared__ float lds[kBT * kMaxRowBlocks * M];
    __shared__ float lds4[4];
    const TileRef t = tiles[blockIdx.x];
    const BatchProblem& p = probs[t.prob];
    if (!lin_active(dyn[t.prob], p.ks)) return;
    float acc[M];
    float d1 = 0.f, d2 = 0.f;
    if (tile_spmv<M, float>(p, t.local, p.J32, p.v[5], lds, acc)) {
        const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
        float* out = which == 0 ? p.v[3] : p.v[4];
        for (int a = 0; a < M; ++a) {
            out[i * M + a] = acc[a];
            if (which == 0) {
                d1 += p.v[1][i * M + a] * acc[a];
            } else {
                d1 += acc[a] * p.v[0][i * M + a];
                d2 += acc[a] * acc[a];
            }
        }
    }
    double part[2];
    part[0] = double(wg_sum<float>(d1, lds4));
    part[1] = double(wg_sum<float>(d2, lds4));
    if (which == 0) put_partials<1>(part, w);
    else put_partials<2>(part, w);
}

template <int M>
__global__ void __launch_bounds__(kBT) k_b_c(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                             const TileRef* __restrict__ tiles, ReduceWork w)
{
    __shared__ float lds4[4];
    const TileRef t = tiles[blockIdx.x];
    const BatchProblem& p = probs[t.prob];
    if (!lin_active(dyn[t.prob], p.ks)) return;
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    float d = 0.f;
    if (i < p.N) {
        const float al = float(p.ks->alpha);
        float rv[M], yv[M];
        for (int a = 0; a < M; ++a) {
            const int64_t k = i * M + a;
            p.v[6][k] += al * p.v[5][k];
            const float ri = p.v[0][k] + (-al) * p.v[3][k];
            p.v[0][k] = ri;
            rv[a] = ri;
            d += ri * ri;
        }
        jac<M>(p.Dinv32, i, rv, yv);
        for (int a = 0; a < M; ++a) p.v[5][i * M + a] = yv[a];
    }
    const double part = double(wg_sum<float>(d, lds4));
    put_partials<1>(&part, w);
}

template <int M>
__global__ void __launch_bounds__(kBT) k_b_e(const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                             const TileRef* __restrict__ tiles, ReduceWork w)
{
    __shared__ float lds4[4];
    const TileRef t = tiles[blockIdx.x];
    const BatchProblem& p = probs[t.prob];
    if (!lin_active(dyn[t.prob], p.ks)) return;
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    float d1 = 0.f, d2 = 0.f;
    if (i < p.N) {
        const float om = float(p.ks->omega);
        for (int a = 0; a < M; ++a) {
            const int64_t k = i * M + a;
            p.v[6][k] += om * p.v[5][k];
            const float ri = p.v[0][k] + (-om) * p.v[4][k];
            p.v[0][k] = ri;
            d1 += ri * ri;
            d2 += p.v[1][k] * ri;
        }
    }
    double part[2];
    part[0] = double(wg_sum<float>(d1, lds4));
    part[1] = double(wg_sum<float>(d2, lds4));
    put_partials<2>(part, w);
}

// one wave per problem of the kind: sum (or max) that problem's tile partials in tile order,
// then the scalar step; NORM / MAX write out[prob]
__global__ void __launch_bounds__(64) k_b_finalize(int op, const BatchProblem* __restrict__ probs, const BatchDyn* __restrict__ dyn,
                                                   const int32_t* __restrict__ plist, int32_t flag, ReduceWork w, double* out,
                                                   double reduction, int32_t maxit)
{
    const int32_t pi = plist[blockIdx.x];
    if (!(dyn[pi].flags & flag)) return;
    const BatchProblem& p = probs[pi];
    const bool usesKs = op != FIN_NORM && op != FIN_MAX;
    if (usesKs && op != FIN_INIT && p.ks->done) return;
    const int nv = (op == FIN_OMEGA || op == FIN_HALF2) ? 2 : 1;
    double tot[2] = {0.0, 0.0};
    for (int k = 0; k < nv; ++k) {
        double acc = 0.0;
        for (int32_t b = threadIdx.x; b < p.ntiles; b += 64) {
            const double v = w.partial[size_t(k) * w.cap + p.tile0 + b];
            acc = op == FIN_MAX ? __builtin_fmax(acc, v) : acc + v;
        }
        if (op == FIN_MAX) {
            for (int o = 32; o > 0; o >>= 1) acc = __builtin_fmax(acc, __shfl_xor(acc, o, 64));
        } else {
            acc = wave_sum<double>(acc);
        }
        tot[k] = acc;
    }
    if (threadIdx.x != 0) return;
    fin_apply(op, tot, p.ks, out + pi, reduction, maxit);
}

} // namespace

} // namespace dmx

// =============================================================================================
// host executor
// =============================================================================================
using dmx::BatchDyn;
using dmx::BatchProblem;
using dmx::TileRef;

struct dmx_batch {
    struct Kind {
        int dim = 0, m = 0;
        int64_t base = 0;                 // first partial slot of this kind's tiles
        std::vector<int32_t> probs;       // global problem indices
        std::vector<TileRef> tiles;
        TileRef* d_tiles = nullptr;
        int32_t* d_plist = nullptr;
    };
    struct NewtonState {
        double t = 0, tEnd = 0, dtNom = 1, dt = 1;
        int steps = 0, retries = 0;
        double shift = 0, lastShift = 0;
        double reduction = 0;             // linear residual reduction (<= 0: the params')
        bool linearPde = false, instationary = true;
        bool finished = false, failed = false, active = false;
        dmx_newton_info info{};
    };

    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<dmx_context*> ctx;
    std::vector<Kind> kinds;
    std::vector<BatchProblem> h_probs;
    std::vector<BatchDyn> h_dyn;
    std::vector<NewtonState> st;
    BatchProblem* d_probs = nullptr;
    BatchDyn* d_dyn = nullptr;
    KrylovScalars* d_ks = nullptr;
    KrylovScalars* h_ks = nullptr;   // pinned
    double* d_out = nullptr;
    double* h_out = nullptr;         // pinned
    int32_t* d_fail = nullptr;
    dmx::ReduceWork rw{};
    std::string err;
    bool tablesDirty = true;
    int64_t capProbs = 0;
};

namespace {

#define B_HIP(call)                                                                  \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            b->err = std::string(#call) + ": " + hipGetErrorString(e_);              \
            return DMX_ERR_DEVICE;                                                   \
        }                                                                            \
    } while (0)
#define B_TRY(expr)                       \
    do {                                  \
        int rc_ = (expr);                 \
        if (rc_ != DMX_OK) return rc_;    \
    } while (0)

template <typename T>
int balloc(dmx_batch* b, T** p, int64_t count)
{
    if (count <= 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * size_t(count));
    if (e != hipSuccess) {
        b->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        *p = nullptr;
        return DMX_ERR_NOMEM;
    }
    return DMX_OK;
}

void free_kind(dmx_batch::Kind& k)
{
    if (k.d_tiles) (void)hipFree(k.d_tiles);
    if (k.d_plist) (void)hipFree(k.d_plist);
    k.d_tiles = nullptr;
    k.d_plist = nullptr;
}

// make sure the context has the arrays of the mixed-precision block-Jacobi solve
int prepare_context(dmx_batch* b, dmx_context* c)
{
    const int mm = c->m * c->m;
    auto need = [&](auto** p, int64_t cnt) -> int {
        if (*p) return DMX_OK;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(**p) * size_t(std::max<int64_t>(cnt, 1)));
        if (e != hipSuccess) { b->err = std::string("hipMalloc: ") + hipGetErrorString(e); return DMX_ERR_NOMEM; }
        return DMX_OK;
    };
    B_TRY(need(&c->J32, c->nnzb * mm));
    B_TRY(need(&c->Dinv32, c->N * mm));
    B_TRY(need(&c->ir_r, c->n));
    for (int i = 0; i < 7; ++i) B_TRY(need(&c->kv32[i], c->n));
    return DMX_OK;
}

BatchProblem describe(dmx_batch* b, int pi)
{
    dmx_context* c = b->ctx[pi];
    BatchProblem p{};
    p.g = c->g;
    p.ph = c->ph;
    AsmPtrs& P = p.P;
    P.u = c->u;
    P.uprev = c->havePrev ? c->uprev : nullptr;
    P.K = c->K;
    P.bc_type = c->bct;
    P.bc_value = c->bcv;
    P.ghost_p = c->ghost_p;
    P.ghost_up = c->ghost_up;
    P.vv_p = c->vvp;
    P.vv_up = c->vvu;
    P.rowptr = c->rowptr;
    P.r = c->r;
    P.J = c->J;
    P.J32 = c->J32;
    P.Dinv32 = c->Dinv32;
    P.fail = b->d_fail;
    P.dt = c->dt;
    P.rowBase = 0;
    p.vvp = c->vvp;
    p.vvu = c->vvu;
    p.rowptr = c->rowptr;
    p.colidx = c->colidx;
    p.J32 = c->J32;
    p.Dinv32 = c->Dinv32;
    p.J = c->J;
    for (int i = 0; i < 7; ++i) p.v[i] = c->kv32[i];
    p.r = c->r;
    p.du = c->du;
    p.ir = c->ir_r;
    p.u = c->u;
    p.u_last = c->u_last;
    p.uprev = c->uprev;
    p.ks = b->d_ks + pi;
    p.N = c->N;
    p.n = c->n;
    return p;
}

int rebuild_tables(dmx_batch* b)
{
    const int P = int(b->ctx.size());
    for (auto& k : b->kinds) free_kind(k);
    b->kinds.clear();
    b->h_probs.assign(P, BatchProblem{});
    for (int pi = 0; pi < P; ++pi) {
        dmx_context* c = b->ctx[pi];
        B_TRY(prepare_context(b, c));
        int ki = -1;
        for (size_t k = 0; k < b->kinds.size(); ++k)
            if (b->kinds[k].dim == c->g.dim && b->kinds[k].m == c->m) ki = int(k);
        if (ki < 0) {
            b->kinds.push_back(dmx_batch::Kind{});
            ki = int(b->kinds.size()) - 1;
            b->kinds[ki].dim = c->g.dim;
            b->kinds[ki].m = c->m;
        }
        auto& K = b->kinds[ki];
        BatchProblem p = describe(b, pi);
        p.tile0 = int32_t(K.tiles.size());  // kind-local here, made global below
        p.ntiles = int32_t((c->N + dmx::kBT - 1) / dmx::kBT);
        for (int32_t t = 0; t < p.ntiles; ++t) K.tiles.push_back(TileRef{pi, t});
        K.probs.push_back(pi);
        b->h_probs[pi] = p;
    }
    int64_t totalTiles = 0;
    for (auto& K : b->kinds) {
        K.base = totalTiles;
        for (int32_t pi : K.probs) b->h_probs[pi].tile0 += int32_t(K.base);
        totalTiles += int64_t(K.tiles.size());
        B_TRY(balloc(b, &K.d_tiles, int64_t(K.tiles.size())));
        B_TRY(balloc(b, &K.d_plist, int64_t(K.probs.size())));
        B_HIP(hipMemcpyAsync(K.d_tiles, K.tiles.data(), sizeof(TileRef) * K.tiles.size(), hipMemcpyHostToDevice, b->stream));
        B_HIP(hipMemcpyAsync(K.d_plist, K.probs.data(), sizeof(int32_t) * K.probs.size(), hipMemcpyHostToDevice, b->stream));
    }
    if (P > b->capProbs) {
        if (b->d_probs) (void)hipFree(b->d_probs);
        if (b->d_dyn) (void)hipFree(b->d_dyn);
        if (b->d_ks) (void)hipFree(b->d_ks);
        if (b->d_out) (void)hipFree(b->d_out);
        if (b->h_ks) (void)hipHostFree(b->h_ks);
        if (b->h_out) (void)hipHostFree(b->h_out);
        b->capProbs = std::max<int64_t>(P, 2 * b->capProbs);
        B_TRY(balloc(b, &b->d_probs, b->capProbs));
        B_TRY(balloc(b, &b->d_dyn, b->capProbs));
        B_TRY(balloc(b, &b->d_ks, b->capProbs));
        B_TRY(balloc(b, &b->d_out, b->capProbs));
        B_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_ks), sizeof(KrylovScalars) * b->capProbs, hipHostMallocDefault));
        B_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_out), sizeof(double) * b->capProbs, hipHostMallocDefault));
        // the KrylovScalars pointers moved
        for (int pi = 0; pi < P; ++pi) b->h_probs[pi].ks = b->d_ks + pi;
        for (int pi = 0; pi < P; ++pi) b->h_probs[pi].P.fail = b->d_fail;
    }
    if (b->rw.cap < totalTiles) {
        if (b->rw.partial) (void)hipFree(b->rw.partial);
        b->rw.cap = totalTiles;
        B_TRY(balloc(b, &b->rw.partial, 2 * b->rw.cap));
    }
    if (P > 0)
        B_HIP(hipMemcpyAsync(b->d_probs, b->h_probs.data(), sizeof(BatchProblem) * P, hipMemcpyHostToDevice, b->stream));
    b->h_dyn.resize(P);
    b->tablesDirty = false;
    return DMX_OK;
}

int upload_dyn(dmx_batch* b)
{
    if (b->h_dyn.empty()) return DMX_OK;
    B_HIP(hipMemcpyAsync(b->d_dyn, b->h_dyn.data(), sizeof(BatchDyn) * b->h_dyn.size(), hipMemcpyHostToDevice, b->stream));
    return DMX_OK;
}

// launch `f(kind)` for every kind with tiles
template <class F>
int each_kind(dmx_batch* b, F&& f)
{
    for (auto& K : b->kinds)
        if (!K.tiles.empty()) {
            hipError_t e = f(K);
            if (e != hipSuccess) {
                b->err = std::string("batched kernel: ") + hipGetErrorString(e);
                return DMX_ERR_DEVICE;
            }
        }
    return DMX_OK;
}

#define KLAUNCH(kernel, K, ...)                                                                       \
    [&]() -> hipError_t {                                                                             \
        const dim3 g_(unsigned(K.tiles.size())), t_(dmx::kBT);                                        \
        if (K.m == 1) kernel<1><<<g_, t_, 0, b->stream>>>(__VA_ARGS__);                               \
        else kernel<2><<<g_, t_, 0, b->stream>>>(__VA_ARGS__);                                        \
        return hipGetLastError();                                                                     \
    }()

// partial slots of one kind's tiles
dmx::ReduceWork kind_rw(const dmx_batch* b, const dmx_batch::Kind& K)
{
    return dmx::ReduceWork{b->rw.partial + K.base, b->rw.cap};
}

hipError_t finalize(dmx_batch* b, dmx_batch::Kind& K, int op, int32_t flag, double reduction, int32_t maxit)
{
    hipLaunchKernelGGL(dmx::k_b_finalize, dim3(unsigned(K.probs.size())), dim3(64), 0, b->stream, op, b->d_probs, b->d_dyn,
                       K.d_plist, flag, b->rw, b->d_out, reduction, maxit);
    return hipGetLastError();
}

int read_out(dmx_batch* b)
{
    B_HIP(hipMemcpyAsync(b->h_out, b->d_out, sizeof(double) * b->ctx.size(), hipMemcpyDeviceToHost, b->stream));
    B_HIP(hipStreamSynchronize(b->stream));
    return DMX_OK;
}

int assemble(dmx_batch* b)
{
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_state, K, b->d_probs, b->d_dyn, K.d_tiles); }));
    return each_kind(b, [&](auto& K) -> hipError_t {
        const dim3 g(unsigned(K.tiles.size())), t(dmx::kBT);
        if (K.dim == 2 && K.m == 1) hipLaunchKernelGGL((dmx::k_b_assemble<2, 1>), g, t, 0, b->stream, b->d_probs, b->d_dyn, K.d_tiles);
        else if (K.dim == 2 && K.m == 2) hipLaunchKernelGGL((dmx::k_b_assemble<2, 2>), g, t, 0, b->stream, b->d_probs, b->d_dyn, K.d_tiles);
        else if (K.dim == 3 && K.m == 1) hipLaunchKernelGGL((dmx::k_b_assemble<3, 1>), g, t, 0, b->stream, b->d_probs, b->d_dyn, K.d_tiles);
        else hipLaunchKernelGGL((dmx::k_b_assemble<3, 2>), g, t, 0, b->stream, b->d_probs, b->d_dyn, K.d_tiles);
        return hipGetLastError();
    });
}

// FP32 block-Jacobi BiCGSTAB inside FP64 refinement for every problem with F_NEWTON set.
// ok[pi] = linear solve converged.
int linear_solve(dmx_batch* b, const dmx_linear_params& lp, std::vector<char>& ok)
{
    const int P = int(b->ctx.size());
    auto& dyn = b->h_dyn;
    // ||r||
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_norm2, K, b->d_probs, b->d_dyn, K.d_tiles, 0, dmx::F_NEWTON, kind_rw(b, K)); }));
    B_TRY(each_kind(b, [&](auto& K) { return finalize(b, K, dmx::FIN_NORM, dmx::F_NEWTON, 0.0, 0); }));
    B_TRY(read_out(b));
    std::vector<double> normb(P, 0.0), res(P, 0.0);
    std::vector<char> conv(P, 0), broken(P, 0);
    for (int pi = 0; pi < P; ++pi) {
        if (!(dyn[pi].flags & dmx::F_NEWTON)) continue;
        normb[pi] = res[pi] = b->h_out[pi];
        conv[pi] = normb[pi] == 0.0;
        if (!(normb[pi] == normb[pi])) broken[pi] = 1;
    }
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_vec, K, b->d_probs, b->d_dyn, K.d_tiles, 0, dmx::F_NEWTON); }));
    for (int k = 0; k < lp.max_refinements; ++k) {
        int nlin = 0;
        for (int pi = 0; pi < P; ++pi) {
            const bool lin = (dyn[pi].flags & dmx::F_NEWTON) && !conv[pi] && !broken[pi];
            dyn[pi].flags = (dyn[pi].flags & ~dmx::F_LIN) | (lin ? dmx::F_LIN : 0);
            dyn[pi].scale = lin ? 1.0 / res[pi] : 0.0;
            dyn[pi].corr = lin ? res[pi] : 0.0;
            nlin += lin;
        }
        if (!nlin) break;
        B_TRY(upload_dyn(b));
        B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_init, K, b->d_probs, b->d_dyn, K.d_tiles, kind_rw(b, K)); }));
        B_TRY(each_kind(b, [&](auto& K) { return finalize(b, K, dmx::FIN_INIT, dmx::F_LIN, lp.inner_reduction, lp.max_iter); }));
        const int G = 8;
        for (int it = 0; it < lp.max_iter; it += G) {
            for (int g = 0; g < G; ++g) {
                B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_a, K, b->d_probs, b->d_dyn, K.d_tiles); }));
                B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_spmv, K, b->d_probs, b->d_dyn, K.d_tiles, kind_rw(b, K), 0); }));
                B_TRY(each_kind(b, [&](auto& K) { return finalize(b, K, dmx::FIN_H, dmx::F_LIN, 0.0, lp.max_iter); }));
                B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_c, K, b->d_probs, b->d_dyn, K.d_tiles, kind_rw(b, K)); }));
                B_TRY(each_kind(b, [&](auto& K) { return finalize(b, K, dmx::FIN_HALF1, dmx::F_LIN, 0.0, lp.max_iter); }));
                B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_spmv, K, b->d_probs, b->d_dyn, K.d_tiles, kind_rw(b, K), 1); }));
                B_TRY(each_kind(b, [&](auto& K) { return finalize(b, K, dmx::FIN_OMEGA, dmx::F_LIN, 0.0, lp.max_iter); }));
                B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_e, K, b->d_probs, b->d_dyn, K.d_tiles, kind_rw(b, K)); }));
                B_TRY(each_kind(b, [&](auto& K) { return finalize(b, K, dmx::FIN_HALF2, dmx::F_LIN, 0.0, lp.max_iter); }));
            }
            B_HIP(hipMemcpyAsync(b->h_ks, b->d_ks, sizeof(KrylovScalars) * P, hipMemcpyDeviceToHost, b->stream));
            B_HIP(hipStreamSynchronize(b->stream));
            bool all = true;
            for (int pi = 0; pi < P; ++pi)
                if ((dyn[pi].flags & dmx::F_LIN) && !b->h_ks[pi].done) all = false;
            if (all) break;
        }
        B_HIP(hipMemcpyAsync(b->h_ks, b->d_ks, sizeof(KrylovScalars) * P, hipMemcpyDeviceToHost, b->stream));
        B_HIP(hipStreamSynchronize(b->stream));
        for (int pi = 0; pi < P; ++pi) {
            if (!(dyn[pi].flags & dmx::F_LIN)) continue;
            const KrylovScalars& ks = b->h_ks[pi];
            b->st[pi].info.linear_iterations += (ks.half_steps + 1) / 2;
            b->st[pi].info.refinements += 1;
            if (ks.status != 0 && ks.half_steps == 0) {  // no progress possible
                broken[pi] = 1;
                dyn[pi].flags &= ~dmx::F_LIN;
            }
        }
        B_TRY(upload_dyn(b));
        B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_correct, K, b->d_probs, b->d_dyn, K.d_tiles); }));
        B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_residual, K, b->d_probs, b->d_dyn, K.d_tiles, kind_rw(b, K)); }));
        B_TRY(each_kind(b, [&](auto& K) { return finalize(b, K, dmx::FIN_NORM, dmx::F_LIN, 0.0, 0); }));
        B_TRY(read_out(b));
        for (int pi = 0; pi < P; ++pi) {
            if (!(dyn[pi].flags & dmx::F_LIN)) continue;
            res[pi] = b->h_out[pi];
            if (!(res[pi] == res[pi])) broken[pi] = 1;
            else if (res[pi] <= (b->st[pi].reduction > 0 ? b->st[pi].reduction : lp.reduction) * normb[pi]) conv[pi] = 1;
        }
    }
    ok.assign(P, 0);
    for (int pi = 0; pi < P; ++pi) {
        dyn[pi].flags &= ~dmx::F_LIN;
        ok[pi] = (dyn[pi].flags & dmx::F_NEWTON) && conv[pi] && !broken[pi];
    }
    return DMX_OK;
}

int vec_op(dmx_batch* b, int op, int32_t flag)
{
    B_TRY(upload_dyn(b));
    return each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_vec, K, b->d_probs, b->d_dyn, K.d_tiles, op, flag); });
}

} // namespace

// One Newton iteration of every active problem; time-step bookkeeping per problem
// (NewtonSolver::solve(vars, timeLoop), newtonsolver.hh:309-355, 428-446, 976-1072).
static int batch_phase(dmx_batch* b, const dmx_newton_params& np, int* nactive)
{
    const int P = int(b->ctx.size());
    if (b->tablesDirty) B_TRY(rebuild_tables(b));
    auto& dyn = b->h_dyn;
    int act = 0;
    for (int pi = 0; pi < P; ++pi) {
        auto& s = b->st[pi];
        const bool on = s.active && !s.finished && !s.failed;
        dyn[pi].flags = on ? dmx::F_NEWTON : 0;
        dyn[pi].dt = s.dt;
        if (on) {
            s.lastShift = s.shift;
            ++act;
        }
    }
    *nactive = act;
    if (!act) return DMX_OK;
    auto t0 = std::chrono::steady_clock::now();
    B_TRY(upload_dyn(b));
    B_TRY(assemble(b));
    auto t1 = std::chrono::steady_clock::now();
    std::vector<char> ok;
    B_TRY(linear_solve(b, np.linear, ok));
    auto t2 = std::chrono::steady_clock::now();
    // problems whose linear solve failed: Newton failure (NumericalProblem) -> no update
    for (int pi = 0; pi < P; ++pi)
        if ((dyn[pi].flags & dmx::F_NEWTON) && !ok[pi]) dyn[pi].flags = 0;
    B_TRY(upload_dyn(b));
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_b_update, K, b->d_probs, b->d_dyn, K.d_tiles, kind_rw(b, K)); }));
    B_TRY(each_kind(b, [&](auto& K) { return finalize(b, K, dmx::FIN_MAX, dmx::F_NEWTON, 0.0, 0); }));
    B_TRY(read_out(b));
    auto t3 = std::chrono::steady_clock::now();
    const double ta = std::chrono::duration<double>(t1 - t0).count(), ts = std::chrono::duration<double>(t2 - t1).count(),
                 tu = std::chrono::duration<double>(t3 - t2).count();
    // per-problem Newton control
    for (int pi = 0; pi < P; ++pi) dyn[pi].flags = 0;
    bool anyAdvance = false, anyReset = false;
    for (int pi = 0; pi < P; ++pi) {
        auto& s = b->st[pi];
        if (!s.active || s.finished || s.failed) continue;
        s.info.t_assemble += ta / act;
        s.info.t_solve += ts / act;
        s.info.t_update += tu / act;
        bool success = false, failure = false;
        if (!ok[pi]) {
            failure = true;
        } else {
            s.shift = b->h_out[pi];
            ++s.steps;
            const bool converged = s.shift <= np.max_rel_shift;
            if (s.linearPde) {
                success = true;
            } else {
                bool proceed;
                if (s.steps < np.min_steps) proceed = true;
                else if (converged) proceed = false;
                else if (s.steps >= np.max_steps) proceed = s.shift * 4.0 < s.lastShift;
                else proceed = true;
                if (!proceed) {
                    if (converged) success = true;
                    else failure = true;
                }
            }
        }
        if (success) {
            s.info.newton_steps += s.steps;
            s.info.last_shift = s.shift;
            s.steps = 0;
            s.retries = 0;
            s.t += s.dt;
            s.info.time_steps++;
            if (s.instationary) {
                dyn[pi].flags |= dmx::F_ADVANCE;
                anyAdvance = true;
            }
            if (s.linearPde || s.t >= s.tEnd) {
                s.finished = true;
                s.info.converged = 1;
            } else {
                s.dt = std::min(s.dtNom, s.tEnd - s.t);
            }
        } else if (failure) {
            s.info.wasted_steps += s.steps;
            s.steps = 0;
            if (s.instationary && s.retries < np.max_dt_divisions) {
                s.retries++;
                s.info.dt_retries++;
                s.dt *= np.retry_factor;
                dyn[pi].flags |= dmx::F_RESET;
                anyReset = true;
            } else {
                s.failed = true;
                s.info.converged = 0;
            }
        }
    }
    if (anyAdvance) B_TRY(vec_op(b, 1, dmx::F_ADVANCE));
    if (anyReset) B_TRY(vec_op(b, 2, dmx::F_RESET));
    // the contexts' solutions are complete when the call returns (their own streams may read them)
    B_HIP(hipStreamSynchronize(b->stream));
    return DMX_OK;
}

extern "C" {

int dmx_batch_create(int32_t device, dmx_batch** out)
{
    if (!out) return DMX_ERR_INVALID;
    auto* b = new dmx_batch();
    b->device = device;
    auto fail = [&](int rc) {
        std::fprintf(stderr, "dmx_batch_create: %s\n", b->err.c_str());
        delete b;
        *out = nullptr;
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        b->err = "cannot create stream";
        return fail(DMX_ERR_DEVICE);
    }
    if (balloc(b, &b->d_fail, 1) != DMX_OK) return fail(DMX_ERR_NOMEM);
    *out = b;
    return DMX_OK;
}

int dmx_batch_destroy(dmx_batch* b)
{
    if (!b) return DMX_ERR_INVALID;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& k : b->kinds) free_kind(k);
    for (void* p : {(void*)b->d_probs, (void*)b->d_dyn, (void*)b->d_ks, (void*)b->d_out, (void*)b->d_fail, (void*)b->rw.partial})
        if (p) (void)hipFree(p);
    if (b->h_ks) (void)hipHostFree(b->h_ks);
    if (b->h_out) (void)hipHostFree(b->h_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return DMX_OK;
}

const char* dmx_batch_last_error(const dmx_batch* b) { return b ? b->err.c_str() : "null batch"; }

int dmx_batch_add(dmx_batch* b, dmx_context* c, const dmx_batch_item* item)
{
    if (!b || !c || !item || !(item->dt > 0.0) || item->nsteps < 0) return DMX_ERR_INVALID;
    if (c->device != b->device || !c->haveState) {
        b->err = "context on another device or without a solution";
        return DMX_ERR_INVALID;
    }
    for (auto* x : b->ctx)
        if (x == c) { b->err = "context already in the batch"; return DMX_ERR_INVALID; }
    B_HIP(hipSetDevice(b->device));
    // the context's own stream may still hold work (uploads)
    B_HIP(hipStreamSynchronize(c->stream));
    dmx_batch::NewtonState s;
    s.linearPde = item->linear_pde != 0;
    s.instationary = !s.linearPde;
    if (s.instationary && !c->havePrev) {
        // uprev <- u (first time level)
        hipError_t e = hipMemcpyAsync(c->uprev, c->u, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) { b->err = "uprev init"; return DMX_ERR_DEVICE; }
        c->havePrev = true;
    }
    if (!s.instationary) c->havePrev = false;
    s.reduction = item->reduction;
    s.dtNom = item->dt;
    s.tEnd = s.linearPde ? item->dt : item->nsteps * item->dt;
    s.dt = std::min(s.dtNom, s.tEnd);
    s.active = true;
    c->jacValid = false;  // the batch overwrites the context's Jacobian and vectors
    b->ctx.push_back(c);
    b->st.push_back(s);
    b->tablesDirty = true;
    return DMX_OK;
}

int dmx_batch_remove(dmx_batch* b, dmx_context* c)
{
    if (!b || !c) return DMX_ERR_INVALID;
    for (size_t i = 0; i < b->ctx.size(); ++i)
        if (b->ctx[i] == c) {
            B_HIP(hipStreamSynchronize(b->stream));
            b->ctx.erase(b->ctx.begin() + i);
            b->st.erase(b->st.begin() + i);
            b->tablesDirty = true;
            return DMX_OK;
        }
    b->err = "context not in the batch";
    return DMX_ERR_INVALID;
}

int dmx_batch_size(const dmx_batch* b) { return b ? int(b->ctx.size()) : -1; }

int dmx_batch_step(dmx_batch* b, const dmx_newton_params* p, int32_t* nactive)
{
    if (!b || !p) return DMX_ERR_INVALID;
    B_HIP(hipSetDevice(b->device));
    int act = 0;
    const int rc = batch_phase(b, *p, &act);
    if (nactive) *nactive = act;
    return rc;
}

int dmx_batch_status(const dmx_batch* b, dmx_context* c, int32_t* finished, dmx_newton_info* info)
{
    if (!b || !c) return DMX_ERR_INVALID;
    for (size_t i = 0; i < b->ctx.size(); ++i)
        if (b->ctx[i] == c) {
            const auto& s = b->st[i];
            if (finished) *finished = s.finished ? 1 : (s.failed ? -1 : 0);
            if (info) *info = s.info;
            return DMX_OK;
        }
    return DMX_ERR_INVALID;
}

int dmx_batch_run(dmx_batch* b, const dmx_newton_params* p)
{
    if (!b || !p) return DMX_ERR_INVALID;
    B_HIP(hipSetDevice(b->device));
    for (;;) {
        int act = 0;
        B_TRY(batch_phase(b, *p, &act));
        if (!act) break;
    }
    B_HIP(hipStreamSynchronize(b->stream));
    return DMX_OK;
}

} // extern "C"
