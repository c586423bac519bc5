// Batched executor with iteration-level continuous batching: many independent problems
// (contexts) advanced together on one GPU, each at its own pace.
//
// Layout: every problem keeps its device arrays in its own context; the batch holds a device
// table of per-problem descriptors (BatchProblem) and, per kind (dim, unknowns per cell), a tile
// table: tile = 256 consecutive cells / block rows of one problem, tiles never straddle
// problems. Every batched kernel runs one workgroup per tile and looks its problem up in the
// table, so the block-diagonal system of all problems is concatenated logically without
// copying.
//
// Control on the device: each problem carries a state machine (ProbState) of the Newton
// iteration (NewtonSolver::solveImpl_, newtonsolver.hh:976-1072; newtonProceed :428-446;
// time-step retry of solve(vars, timeLoop) :309-355) and of the FP64 iterative refinement
// around its FP32 BiCGSTAB. A "tick" is a fixed sequence of launches, one per phase:
//   update -> Newton control -> time level (advance / reset) -> assembly -> ||r|| -> control
//   -> Krylov init -> control -> kTickIters BiCGSTAB iterations -> control -> correction +
//   FP64 refinement residual -> control
// Every kernel acts only on the problems in its phase, and the control kernels (one wave per
// problem) move a problem to its next phase, so within one tick a problem can finish its
// linear solve, update, re-assemble and start the next solve while others keep iterating.
// The tick does not depend on the problems' states, so it is captured once per batch
// composition as a HIP graph; the host only launches ticks, reads the states, and admits or
// retires problems between ticks.
//
// Linear solves: FP64 iterative refinement (residual ir = r - J du with the FP64 BCRS Jacobian)
// around BiCGSTAB in FP32 on the block-Jacobi left-scaled operator D^-1 J, which the assembly
// writes in the structured-stencil layout of stencil.hh (no column indices; the scaled diagonal
// is the identity, so the preconditioner costs nothing inside the iteration).
//
// Determinism: a reduction over a problem sums that problem's tile partials in tile order, and
// tiles start at the problem's first row, so a problem's results are bitwise the same alone
// and in any mix, whenever other problems join or leave. Workgroups are mapped to tiles so that
// each XCD (own L2) works on a contiguous range of tiles, which keeps stencil neighbours local.
#include "context.hh"
#include "krylov_ops.hh"
#include "stencil.hh"
#include "mg_sweeps.hh"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace dmx {

namespace {

constexpr int kBT = 256;           // cells per tile = threads per workgroup
constexpr int kMaxRowBlocks = 7;
constexpr int kTickIters = 8;      // BiCGSTAB iterations per tick
constexpr int kXcds = 8;

enum Phase : int32_t {
    PH_IDLE = 0,      // finished, failed
    PH_ASM = 1,       // assemble residual and Jacobian at u
    PH_ADVANCE = 2,   // uprev <- u, then assemble (next time step)
    PH_RESET = 3,     // u <- uprev, then assemble (retry with a smaller dt)
    PH_INIT = 4,      // start an FP32 solve of the refinement residual
    PH_KRYLOV = 5,    // BiCGSTAB iterations
    PH_CORRECT = 6,   // du += correction, FP64 refinement residual
    PH_UPDATE = 7,    // u <- u - du, maximum relative shift
    PH_RESTART = 8    // restart the linear solve (du = 0) in FP64
};

struct ProbState {
    int32_t phase;
    int32_t status;          // 0 running, 1 finished, -1 failed (time-step divisions exhausted)
    int32_t steps;           // Newton iterations of the current attempt
    int32_t retries;         // dt divisions of the current time step
    int32_t ref;             // refinement index of the current linear solve
    int32_t linearPde;       // one stationary step
    int32_t prec64;          // linear solves in FP64 (the FP32 refinement stalled)
    int32_t stalls;          // consecutive FP32 refinements that reduced the residual less than 4x
    int32_t newtonSteps, wastedSteps, linIts, refinements, timeSteps, dtRetries;
    double t, tEnd, dt, dtNom;
    double shift, lastShift;
    double normb, res;       // ||r|| of the Newton residual, current refinement residual
    double resPrev;          // refinement residual before the last correction
    double reduction;        // linear residual reduction of this problem
    double scale, corr;      // FP32 right-hand side scaling (1/res) and correction scaling (res)
};

struct BatchParams {
    double maxRelShift;
    int32_t minSteps, maxSteps, maxDiv, maxIter, maxRef, targetSteps;
    double retryFactor;
    double innerReduction;
    int32_t adaptive;
    int32_t mg;               // 1: aggregation multigrid preconditions the FP32 BiCGSTAB
    double maxDt;             // adaptive dt cap (<= 0: none)
};

// NewtonSolver::suggestTimeStepSize (newtonsolver.hh:784-798)
__device__ __forceinline__ double suggest_time_step(double oldDt, int32_t numSteps, int32_t targetSteps)
{
    if (numSteps > targetSteps) return oldDt / (1.0 + double(numSteps - targetSteps) / targetSteps);
    return oldDt * (1.0 + (double(targetSteps - numSteps) / targetSteps) / 1.2);
}

struct BatchProblem {
    GridDesc g;
    Phys ph;
    AsmPtrs P;                 // assembly pointers; P.S32 and P.Dinv32 are set
    double* vvp;
    double* vvu;
    double* vvd;      // 2p analytic Jacobian: saturation derivatives (null otherwise)
    const int32_t* rowptr;
    const int32_t* colidx;
    float* S;                  // FP32 off-diagonal blocks (stencil layout), scaled by D^-1 in place
    const float* Dinv32;
    const double* Dinv64;
    const double* J;
    float* v32[8];             // FP32 Krylov vectors (aliases of v64's memory): r rt p v t x, then
                               // with multigrid y (preconditioned direction) and the level-0 iterate
    double* v64[7];            // FP64 Krylov vectors
    double* r;                 // Newton residual
    double* du;                // Newton update
    double* ir;                // refinement residual
    double* u;
    double* u_last;
    double* uprev;
    KrylovScalars* ks;
    int64_t N, n;
    int32_t tile0, ntiles;     // this problem's partial slots: tile0 .. tile0 + ntiles - 1
    int32_t nlev;              // multigrid levels (1: no multigrid, the identity preconditioner)
    int32_t mgTail;            // first multigrid level of the one-workgroup tail (mg_tail_level)
    MgLevel lev[kMgMaxLevels]; // level 0 = the scaled operator S
};

struct TileRef {
    int32_t prob;              // problem index
    int32_t local;             // tile index inside the problem
};

// Work lists: after every control stage a plan kernel compacts, per phase group, the problems
// whose phase belongs to the group into segments of consecutive tiles (exclusive prefix sums
// of their tile counts). The phase kernels run a fixed grid that loops over the active tiles
// only, so problems in other phases cost nothing.
enum Group : int32_t { G_UPDATE = 0, G_LEVEL, G_ASM, G_BEGIN, G_INIT, G_KRYLOV, G_CORRECT, kGroups };

struct Work {
    int32_t* count;            // [kGroups] active problems
    int32_t* total;            // [kGroups] active tiles
    int32_t* start;            // [kGroups * cap] first active tile of each active problem
    int32_t* prob;             // [kGroups * cap] the active problems
    int32_t cap;               // problems of the kind
    TileRef* kmap;             // active tile -> tile of the BiCGSTAB group (filled by k_t_expand)
};

// Walks this workgroup's share [u0, u1) of the active tiles of group g: consecutive workgroups
// go round-robin to the XCDs, so XCD k gets a contiguous share. The segment of the first tile is
// found by binary search, later tiles step through the segments in order.
struct TileIter {
    const int32_t* start;
    const int32_t* prob;
    int32_t count, total;
    int64_t u, u1;
    int32_t j, segEnd;

    __device__ __forceinline__ TileIter(const Work& W, int g)
    {
        start = W.start + int64_t(g) * W.cap;
        prob = W.prob + int64_t(g) * W.cap;
        count = W.count[g];
        total = W.total[g];
        const int64_t nb = gridDim.x;
        const int64_t w = (int64_t(blockIdx.x) % kXcds) * (nb / kXcds) + int64_t(blockIdx.x) / kXcds;
        const int64_t chunk = (int64_t(total) + nb - 1) / nb;
        u = w * chunk;
        u1 = u + chunk < total ? u + chunk : total;
        j = 0;
        segEnd = 0;
        if (u < u1) {
            int32_t lo = 0, hi = count - 1;
            while (lo < hi) {
                const int32_t mid = (lo + hi + 1) >> 1;
                if (start[mid] <= u) lo = mid;
                else hi = mid - 1;
            }
            j = lo;
            segEnd = j + 1 < count ? start[j + 1] : total;
        }
    }

    // next active tile -> (problem, tile of the problem) and its partial-sum slot
    __device__ __forceinline__ bool next(const BatchProblem* __restrict__ probs, TileRef& t, int64_t& tix)
    {
        if (u >= u1) return false;
        while (u >= segEnd) {
            ++j;
            segEnd = j + 1 < count ? start[j + 1] : total;
        }
        t.prob = prob[j];
        t.local = int32_t(u - start[j]);
        tix = int64_t(probs[t.prob].tile0) + t.local;
        ++u;
        return true;
    }
};

__device__ __forceinline__ bool in_asm(int32_t ph) { return ph == PH_ASM || ph == PH_ADVANCE || ph == PH_RESET; }

__device__ __forceinline__ bool in_group(int32_t ph, int g)
{
    switch (g) {
    case G_UPDATE: return ph == PH_UPDATE;
    case G_LEVEL: return ph == PH_ADVANCE || ph == PH_RESET;
    case G_ASM: return in_asm(ph);
    case G_BEGIN: return in_asm(ph) || ph == PH_RESTART;
    case G_INIT: return ph == PH_INIT;
    case G_KRYLOV: return ph == PH_KRYLOV;
    default: return ph == PH_CORRECT;
    }
}

// exclusive prefix sum over a 1024-thread workgroup; *total gets the sum
__device__ __forceinline__ int32_t block_exscan(int32_t v, int32_t* lds16, int32_t* total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds16[wave] = x;
    __syncthreads();
    int32_t before = 0, all = 0;
    for (int k = 0; k < 16; ++k) {
        const int32_t wv = lds16[k];
        if (k < wave) before += wv;
        all += wv;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// one workgroup per kind: the work lists of every phase group
__global__ void __launch_bounds__(1024) k_t_plan(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st,
                                                 const int32_t* __restrict__ plist, Work W)
{
    __shared__ int32_t lds16[16];
    for (int g = 0; g < kGroups; ++g) {
        int32_t nprob = 0, ntile = 0;
        for (int32_t c0 = 0; c0 < W.cap; c0 += 1024) {
            const int32_t j = c0 + int32_t(threadIdx.x);
            const int32_t pi = j < W.cap ? plist[j] : 0;
            const bool on = j < W.cap && in_group(st[pi].phase, g);
            int32_t sumProb = 0, sumTile = 0;
            const int32_t ep = block_exscan(on ? 1 : 0, lds16, &sumProb);
            const int32_t et = block_exscan(on ? probs[pi].ntiles : 0, lds16, &sumTile);
            if (on) {
                W.start[int64_t(g) * W.cap + nprob + ep] = ntile + et;
                W.prob[int64_t(g) * W.cap + nprob + ep] = pi;
            }
            nprob += sumProb;
            ntile += sumTile;
        }
        if (threadIdx.x == 0) {
            W.count[g] = nprob;
            W.total[g] = ntile;
        }
    }
}

// one workgroup per problem slot: the BiCGSTAB group's active tiles as a direct map
__global__ void __launch_bounds__(256) k_t_expand(const BatchProblem* __restrict__ probs, Work W)
{
    const int32_t j = blockIdx.x;
    if (j >= W.count[G_KRYLOV]) return;
    const int32_t pi = W.prob[int64_t(G_KRYLOV) * W.cap + j];
    const int32_t s0 = W.start[int64_t(G_KRYLOV) * W.cap + j];
    const int32_t nt = probs[pi].ntiles;
    for (int32_t k = threadIdx.x; k < nt; k += blockDim.x) W.kmap[s0 + k] = TileRef{pi, k};
}

// this workgroup's share [u, u1) of the BiCGSTAB group's active tiles (XCD-contiguous)
__device__ __forceinline__ void krylov_range(const Work& W, int64_t& u, int64_t& u1)
{
    const int64_t total = W.total[G_KRYLOV];
    const int64_t nb = gridDim.x;
    const int64_t w = (int64_t(blockIdx.x) % kXcds) * (nb / kXcds) + int64_t(blockIdx.x) / kXcds;
    const int64_t chunk = (total + nb - 1) / nb;
    u = w * chunk;
    u1 = u + chunk < total ? u + chunk : total;
}

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
__device__ __forceinline__ void put_partials(const double* part, ReduceWork w, int64_t tix)
{
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) w.partial[size_t(k) * w.cap + tix] = part[k];
}

// ---- Newton update and time level ------------------------------------------------------------
// u_last = u; u = u - du; max relative shift partials (newtonsolver.hh:110-116, 543-604)
template <int M>
__device__ __forceinline__ void tile_update(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, ReduceWork w, TileRef t, int64_t tix, double* lds4)
{
    if (st[t.prob].phase != PH_UPDATE) return;
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
    put_partials<1>(&part, w, tix);
}

template <int M>
__global__ void __launch_bounds__(kBT) k_t_update(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W, ReduceWork w)
{
    __shared__ double lds4[4];
    TileIter it(W, G_UPDATE);
    TileRef t;
    int64_t tix;
    while (it.next(probs, t, tix)) tile_update<M>(probs, st, w, t, tix, lds4);
}

// advanceTimeStep (uprev <- u) / resetTimeStep (u <- uprev) (fvassembler.hh:387, fvgridvariables.hh:95-115)
template <int M>
__device__ __forceinline__ void tile_level(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, TileRef t, int64_t tix)
{
    const int32_t ph = st[t.prob].phase;
    if (ph != PH_ADVANCE && ph != PH_RESET) return;
    const BatchProblem& p = probs[t.prob];
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    if (i >= p.N) return;
    for (int a = 0; a < M; ++a) {
        const int64_t k = i * M + a;
        if (ph == PH_ADVANCE) p.uprev[k] = p.u[k];
        else p.u[k] = p.uprev[k];
    }
}

template <int M>
__global__ void __launch_bounds__(kBT) k_t_level(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W)
{
    TileIter it(W, G_LEVEL);
    TileRef t;
    int64_t tix;
    while (it.next(probs, t, tix)) tile_level<M>(probs, st, t, tix);
}

// ---- assembly -------------------------------------------------------------------------------
template <int M>
__device__ __forceinline__ void tile_state(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, TileRef t, int64_t tix)
{
    if (!in_asm(st[t.prob].phase)) return;
    const BatchProblem& p = probs[t.prob];
    const int64_t I = int64_t(t.local) * kBT + threadIdx.x;
    if (I >= p.N) return;
    cell_state<M>(p.ph, p.P.u, p.vvp, p.vvu, p.vvd, p.P.mat, I);
}

template <int M>
__global__ void __launch_bounds__(kBT) k_t_state(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W)
{
    TileIter it(W, G_ASM);
    TileRef t;
    int64_t tix;
    while (it.next(probs, t, tix)) tile_state<M>(probs, st, t, tix);
}

template <int DIM, int M, bool AN>
__device__ __forceinline__ void tile_assemble(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, TileRef t, int64_t tix)
{
    if (!in_asm(st[t.prob].phase)) return;
    const BatchProblem& p = probs[t.prob];
    const int64_t I = int64_t(t.local) * kBT + threadIdx.x;
    if (I >= p.N) return;
    AsmPtrs P = p.P;
    P.dt = st[t.prob].dt;
    assemble_cell<DIM, M, true, AN>(p.g, p.ph, P, I);
}

// AN: the members with hand-coded derivatives (others skip their tiles), else forward differences
template <int DIM, int M, bool AN>
__global__ void __launch_bounds__(kBT) k_t_assemble(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W)
{
    TileIter it(W, G_ASM);
    TileRef t;
    int64_t tix;
    while (it.next(probs, t, tix))
        if ((probs[t.prob].P.analytic != 0) == AN) tile_assemble<DIM, M, AN>(probs, st, t, tix);
}

// after the assembly: scale the FP32 operator rows by D^-1; start (or restart) of a linear
// solve: du = 0, ir = r, ||r||^2 partials
template <int M>
__device__ __forceinline__ void tile_begin(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, ReduceWork w, TileRef t, int64_t tix, double* lds4)
{
    const int32_t ph = st[t.prob].phase;
    if (!in_asm(ph) && ph != PH_RESTART) return;
    const BatchProblem& p = probs[t.prob];
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    double acc = 0.0;
    if (i < p.N) {
        if (ph != PH_RESTART) {
            scale_row<M>(p.g, p.S, p.Dinv32, i);
            if (uint16_t* h = p.lev[0].S16)   // the multigrid sweeps' bfloat16 copy of the row
                for (int d = 0; d < 2 * p.g.dim; ++d)
                    for (int q = 0; q < M * M; ++q)
                        h[(int64_t(d) * p.N + i) * (M * M) + q] = mg_to_bf16(p.S[(int64_t(d) * M * M + q) * p.N + i]);
        }
        for (int a = 0; a < M; ++a) {
            const int64_t k = i * M + a;
            const double ri = p.r[k];
            p.du[k] = 0.0;
            p.ir[k] = ri;
            acc += ri * ri;
        }
    }
    const double part = wg_sum<double>(acc, lds4);
    put_partials<1>(&part, w, tix);
}

template <int M>
__global__ void __launch_bounds__(kBT) k_t_begin(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W, ReduceWork w)
{
    __shared__ double lds4[4];
    TileIter it(W, G_BEGIN);
    TileRef t;
    int64_t tix;
    while (it.next(probs, t, tix)) tile_begin<M>(probs, st, w, t, tix, lds4);
}

// ---- FP64 iterative refinement ---------------------------------------------------------------
// tile SpMV of the FP64 BCRS Jacobian (one tile = 256 consecutive rows of one problem)
template <int M>
__device__ __forceinline__ bool tile_spmv64(const BatchProblem& p, int64_t tileLocal, const double* __restrict__ x,
                                            double* lds, double (&acc)[M])
{
    const int64_t r0 = tileLocal * kBT;
    const int64_t r1 = r0 + kBT < p.N ? r0 + kBT : p.N;
    const int32_t k0 = p.rowptr[r0], k1 = p.rowptr[r1];
    for (int32_t k = k0 + threadIdx.x; k < k1; k += kBT) {
        const int64_t j = p.colidx[k];
        const double* B = p.J + int64_t(k) * M * M;
        double Bv[M * M];
        for (int q = 0; q < M * M; ++q) Bv[q] = B[q];
        double xv[M];
        for (int b = 0; b < M; ++b) xv[b] = x[j * M + b];
        for (int a = 0; a < M; ++a) {
            double s = 0.0;
            for (int b = 0; b < M; ++b) s += Bv[a * M + b] * xv[b];
            lds[(k - k0) * M + a] = s;
        }
    }
    __syncthreads();
    const int64_t row = r0 + threadIdx.x;
    for (int a = 0; a < M; ++a) acc[a] = 0.0;
    if (row >= r1) return false;
    const int32_t e = p.rowptr[row + 1];
    for (int32_t k = p.rowptr[row]; k < e; ++k)
        for (int a = 0; a < M; ++a) acc[a] += lds[(k - k0) * M + a];
    return true;
}

// du += corr * x32 (scaled FP32 correction)
template <int M>
__device__ __forceinline__ void tile_correct(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, TileRef t, int64_t tix)
{
    if (st[t.prob].phase != PH_CORRECT) return;
    const double corr = st[t.prob].corr;
    const BatchProblem& p = probs[t.prob];
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    if (i >= p.N) return;
    if (st[t.prob].prec64)
        for (int a = 0; a < M; ++a) p.du[i * M + a] += corr * p.v64[6][i * M + a];
    else
        for (int a = 0; a < M; ++a) p.du[i * M + a] += corr * double(p.v32[5][i * M + a]);
}

template <int M>
__global__ void __launch_bounds__(kBT) k_t_correct(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W)
{
    TileIter it(W, G_CORRECT);
    TileRef t;
    int64_t tix;
    while (it.next(probs, t, tix)) tile_correct<M>(probs, st, t, tix);
}

// ir = r - J du (FP64) and ||ir||^2 partials
template <int M>
__device__ __forceinline__ void tile_residual(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, ReduceWork w, TileRef t, int64_t tix, double* lds, double* lds4)
{
    if (st[t.prob].phase != PH_CORRECT) return;
    const BatchProblem& p = probs[t.prob];
    double acc[M];
    double nr = 0.0;
    if (tile_spmv64<M>(p, t.local, p.du, lds, acc)) {
        const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
        for (int a = 0; a < M; ++a) {
            const double v = p.r[i * M + a] - acc[a];
            p.ir[i * M + a] = v;
            nr += v * v;
        }
    }
    const double part = wg_sum<double>(nr, lds4);
    put_partials<1>(&part, w, tix);
}

template <int M>
__global__ void __launch_bounds__(kBT) k_t_residual(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W, ReduceWork w)
{
    __shared__ double lds[kBT * kMaxRowBlocks * M];
    __shared__ double lds4[4];
    TileIter it(W, G_CORRECT);
    TileRef t;
    int64_t tix;
    while (it.next(probs, t, tix)) tile_residual<M>(probs, st, w, t, tix, lds, lds4);
}

// ---- BiCGSTAB, per problem scalars -----------------------------------------------------------
// FP32 form on the scaled stencil operator A = D^-1 J (the default), vectors v32:
//   0 r, 1 rt, 2 p, 3 v, 4 t, 5 x
// FP64 form (fallback of a problem whose FP32 refinement stalls) on the FP64 BCRS Jacobian with
// block Jacobi as right preconditioner (dune-istl order, y = D^-1 p, v = J y), vectors v64:
//   0 r, 1 rt, 2 p, 3 v, 4 t, 5 y, 6 x
// Both forms share every launch; a workgroup takes the form of its problem.
__device__ __forceinline__ bool krylov_active(const ProbState& s, const KrylovScalars* ks)
{
    return s.phase == PH_KRYLOV && !ks->done;
}

template <int M>
__device__ __forceinline__ void jac64(const double* __restrict__ Dinv, int64_t i, const double (&x)[M], double (&y)[M])
{
    for (int a = 0; a < M; ++a) {
        double s = 0.0;
        for (int b = 0; b < M; ++b) s += Dinv[i * M * M + a * M + b] * x[b];
        y[a] = s;
    }
}

// y = J x for row i of the FP64 BCRS Jacobian (row per lane)
template <int M>
__device__ __forceinline__ void bcrs_row64(const BatchProblem& p, const double* __restrict__ x, int64_t i, double (&y)[M])
{
    for (int a = 0; a < M; ++a) y[a] = 0.0;
    const int32_t e = p.rowptr[i + 1];
    for (int32_t k = p.rowptr[i]; k < e; ++k) {
        const int64_t j = p.colidx[k];
        const double* B = p.J + int64_t(k) * M * M;
        for (int a = 0; a < M; ++a) {
            double s = 0.0;
            for (int b = 0; b < M; ++b) s += B[a * M + b] * x[j * M + b];
            y[a] += s;
        }
    }
}

// FP32: r = rt = D^-1 (scale * ir), x = 0.   FP64: r = rt = scale * ir, x = 0.   ||r||^2 partials
template <int M>
__device__ __forceinline__ void tile_init(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, ReduceWork w, TileRef t, int64_t tix, double* lds4d, float* lds4f)
{
    if (st[t.prob].phase != PH_INIT) return;
    const double scale = st[t.prob].scale;
    const BatchProblem& p = probs[t.prob];
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    double part;
    if (st[t.prob].prec64) {
        double acc = 0.0;
        if (i < p.N) {
            for (int a = 0; a < M; ++a) {
                const int64_t k = i * M + a;
                const double ri = p.ir[k] * scale;
                p.v64[0][k] = ri;
                p.v64[1][k] = ri;
                p.v64[6][k] = 0.0;
                acc += ri * ri;
            }
        }
        part = wg_sum<double>(acc, lds4d);
    } else {
        float acc = 0.f;
        if (i < p.N) {
            float b[M];
            for (int a = 0; a < M; ++a) b[a] = float(p.ir[i * M + a] * scale);
            for (int a = 0; a < M; ++a) {
                float ri = 0.f;
                for (int c = 0; c < M; ++c) ri += p.Dinv32[i * M * M + a * M + c] * b[c];
                const int64_t k = i * M + a;
                p.v32[0][k] = ri;
                p.v32[1][k] = ri;
                p.v32[5][k] = 0.f;
                acc += ri * ri;
            }
        }
        part = double(wg_sum<float>(acc, lds4f));
    }
    put_partials<1>(&part, w, tix);
}

template <int M>
__global__ void __launch_bounds__(kBT) k_t_init(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W, ReduceWork w)
{
    __shared__ double lds4d[4];
    __shared__ float lds4f[4];
    TileIter it(W, G_INIT);
    TileRef t;
    int64_t tix;
    while (it.next(probs, t, tix)) tile_init<M>(probs, st, w, t, tix, lds4d, lds4f);
}

// A: p = r + beta (p - omega v) (first iteration: p = r); FP64 form also y = D^-1 p
template <int M>
__device__ __forceinline__ void tile_a(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, TileRef t, int64_t tix)
{
    const BatchProblem& p = probs[t.prob];
    if (!krylov_active(st[t.prob], p.ks)) return;
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    if (i >= p.N) return;
    const KrylovScalars* ks = p.ks;
    const bool first = ks->half_steps == 0;
    const double beta = (ks->norm == 0.0) ? 0.0 : (ks->rho_new / ks->rho) * (ks->alpha / ks->omega);
    if (st[t.prob].prec64) {
        double pv[M], yv[M];
        for (int a = 0; a < M; ++a) {
            const int64_t k = i * M + a;
            const double ri = p.v64[0][k];
            double q = ri;
            if (!first) {
                q = p.v64[2][k];
                q += -ks->omega * p.v64[3][k];
                q *= beta;
                q += ri;
            }
            p.v64[2][k] = q;
            pv[a] = q;
        }
        jac64<M>(p.Dinv64, i, pv, yv);
        for (int a = 0; a < M; ++a) p.v64[5][i * M + a] = yv[a];
    } else {
        const float om = float(ks->omega), bt = float(beta);
        for (int a = 0; a < M; ++a) {
            const int64_t k = i * M + a;
            const float ri = p.v32[0][k];
            float q = ri;
            if (!first) {
                q = p.v32[2][k];
                q += -om * p.v32[3][k];
                q *= bt;
                q += ri;
            }
            p.v32[2][k] = q;
        }
    }
}


// B (which = 0): v = A p, <rt, v>;  D (which = 1): t = A r, <t, r>, <t, t>
// (FP64 form: v = J y and t = J y with the preconditioned vector y)
template <int M>
__device__ __forceinline__ void tile_spmv(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, ReduceWork w, int which, TileRef t, int64_t tix, double* lds4d, float* lds4f)
{
    const BatchProblem& p = probs[t.prob];
    if (!krylov_active(st[t.prob], p.ks)) return;
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    double part[2];
    if (st[t.prob].prec64) {
        double d1 = 0.0, d2 = 0.0;
        if (i < p.N) {
            double y[M];
            bcrs_row64<M>(p, p.v64[5], i, y);
            double* out = which == 0 ? p.v64[3] : p.v64[4];
            for (int a = 0; a < M; ++a) {
                out[i * M + a] = y[a];
                if (which == 0) {
                    d1 += p.v64[1][i * M + a] * y[a];
                } else {
                    d1 += y[a] * p.v64[0][i * M + a];
                    d2 += y[a] * y[a];
                }
            }
        }
        part[0] = wg_sum<double>(d1, lds4d);
        part[1] = wg_sum<double>(d2, lds4d);
    } else {
        float d1 = 0.f, d2 = 0.f;
        if (i < p.N) {
            const float* x = which == 0 ? p.v32[2] : p.v32[0];
            float y[M];
            // with multigrid the operator acts on the preconditioned vector y = M^-1 x
            stencil_apply<M>(p.g, p.S, p.nlev > 1 ? p.v32[6] : x, i, y);
            float* out = which == 0 ? p.v32[3] : p.v32[4];
            for (int a = 0; a < M; ++a) {
                out[i * M + a] = y[a];
                if (which == 0) {
                    d1 += p.v32[1][i * M + a] * y[a];
                } else {
                    d1 += y[a] * x[i * M + a];
                    d2 += y[a] * y[a];
                }
            }
        }
        part[0] = double(wg_sum<float>(d1, lds4f));
        part[1] = double(wg_sum<float>(d2, lds4f));
    }
    if (which == 0) put_partials<1>(part, w, tix);
    else put_partials<2>(part, w, tix);
}


// C: x += alpha p, r -= alpha v, ||r||^2   (FP64 form: x += alpha y, ..., then y = D^-1 r)
template <int M>
__device__ __forceinline__ void tile_c(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, ReduceWork w, TileRef t, int64_t tix, double* lds4d, float* lds4f)
{
    const BatchProblem& p = probs[t.prob];
    if (!krylov_active(st[t.prob], p.ks)) return;
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    double part;
    if (st[t.prob].prec64) {
        double d = 0.0;
        if (i < p.N) {
            const double al = p.ks->alpha;
            double rv[M], yv[M];
            for (int a = 0; a < M; ++a) {
                const int64_t k = i * M + a;
                p.v64[6][k] += al * p.v64[5][k];
                const double ri = p.v64[0][k] + (-al) * p.v64[3][k];
                p.v64[0][k] = ri;
                rv[a] = ri;
                d += ri * ri;
            }
            jac64<M>(p.Dinv64, i, rv, yv);
            for (int a = 0; a < M; ++a) p.v64[5][i * M + a] = yv[a];
        }
        part = wg_sum<double>(d, lds4d);
    } else {
        float d = 0.f;
        if (i < p.N) {
            const float al = float(p.ks->alpha);
            const float* dir = p.nlev > 1 ? p.v32[6] : p.v32[2];
            for (int a = 0; a < M; ++a) {
                const int64_t k = i * M + a;
                p.v32[5][k] += al * dir[k];
                const float ri = p.v32[0][k] + (-al) * p.v32[3][k];
                p.v32[0][k] = ri;
                d += ri * ri;
            }
        }
        part = double(wg_sum<float>(d, lds4f));
    }
    put_partials<1>(&part, w, tix);
}


// E: x += omega r, r -= omega t, ||r||^2, <rt, r>   (FP64 form: x += omega y)
template <int M>
__device__ __forceinline__ void tile_e(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, ReduceWork w, TileRef t, int64_t tix, double* lds4d, float* lds4f)
{
    const BatchProblem& p = probs[t.prob];
    if (!krylov_active(st[t.prob], p.ks)) return;
    const int64_t i = int64_t(t.local) * kBT + threadIdx.x;
    double part[2];
    if (st[t.prob].prec64) {
        double d1 = 0.0, d2 = 0.0;
        if (i < p.N) {
            const double om = p.ks->omega;
            for (int a = 0; a < M; ++a) {
                const int64_t k = i * M + a;
                p.v64[6][k] += om * p.v64[5][k];
                const double ri = p.v64[0][k] + (-om) * p.v64[4][k];
                p.v64[0][k] = ri;
                d1 += ri * ri;
                d2 += p.v64[1][k] * ri;
            }
        }
        part[0] = wg_sum<double>(d1, lds4d);
        part[1] = wg_sum<double>(d2, lds4d);
    } else {
        float d1 = 0.f, d2 = 0.f;
        if (i < p.N) {
            const float om = float(p.ks->omega);
            const float* dir = p.nlev > 1 ? p.v32[6] : p.v32[0];
            for (int a = 0; a < M; ++a) {
                const int64_t k = i * M + a;
                const float s = p.v32[0][k];
                p.v32[5][k] += om * dir[k];
                const float ri = s + (-om) * p.v32[4][k];
                p.v32[0][k] = ri;
                d1 += ri * ri;
                d2 += p.v32[1][k] * ri;
            }
        }
        part[0] = double(wg_sum<float>(d1, lds4f));
        part[1] = double(wg_sum<float>(d2, lds4f));
    }
    put_partials<2>(part, w, tix);
}


// ---- coarsened FP32 BiCGSTAB kernels ---------------------------------------------------------
// A workgroup takes up to C consecutive tiles of one problem per pass (C, and the grid, are
// chosen per batch by dmx_batch_autotune): the problem's
// descriptor, state and scalars are fetched once, every thread issues the loads of its cell in
// each tile before using them, and the per-tile partial sums are reduced together. Each tile's
// partial is the same sum in the same order as a one-tile workgroup computes.
struct TileRun {
    TileRef t;     // first tile
    int64_t tix;   // its partial-sum slot (the run's slots are consecutive)
    int32_t c;     // tiles in the run
};

// the next run of map entries [u, u1): consecutive tiles of one problem
template <int C>
__device__ __forceinline__ bool next_run(const Work& W, int64_t& u, int64_t u1, const BatchProblem* __restrict__ probs,
                                         TileRun& r)
{
    if (u >= u1) return false;
    r.t = W.kmap[u];
    r.c = 1;
    while (r.c < C && u + r.c < u1 && W.kmap[u + r.c].prob == r.t.prob) ++r.c;
    r.tix = int64_t(probs[r.t.prob].tile0) + r.t.local;
    u += r.c;
    return true;
}

// per-tile sums of NV values per tile over the workgroup (wave shuffles, then the 4 waves in
// fixed order); thread 0 writes slot tix + k of value v at partial[v*cap + tix + k]
template <int NV, int C>
__device__ __forceinline__ void run_partials(float (&acc)[C][NV], int c, float* lds, ReduceWork w, int64_t tix)
{
    const int wave = threadIdx.x >> 6;
    for (int k = 0; k < C; ++k)
        for (int v = 0; v < NV; ++v) {
            const float s = wave_sum<float>(acc[k][v]);
            if ((threadIdx.x & 63) == 0) lds[(k * NV + v) * 4 + wave] = s;
        }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < c; ++k)
            for (int v = 0; v < NV; ++v) {
                const float* q = lds + (k * NV + v) * 4;
                w.partial[size_t(v) * w.cap + tix + k] = double((q[0] + q[1]) + (q[2] + q[3]));
            }
    __syncthreads();
}

template <int M, int C>
__device__ __forceinline__ void run_a32(const BatchProblem& p, const TileRun& r)
{
    const KrylovScalars* ks = p.ks;
    const bool first = ks->half_steps == 0;
    const double beta = (ks->norm == 0.0) ? 0.0 : (ks->rho_new / ks->rho) * (ks->alpha / ks->omega);
    const float om = float(ks->omega), bt = float(beta);
    float rv[C][M], pv[C][M], vv[C][M];
    for (int k = 0; k < C; ++k) {
        const int64_t i = int64_t(r.t.local + k) * kBT + threadIdx.x;
        if (k < r.c && i < p.N)
            for (int a = 0; a < M; ++a) {
                rv[k][a] = p.v32[0][i * M + a];
                if (!first) {
                    pv[k][a] = p.v32[2][i * M + a];
                    vv[k][a] = p.v32[3][i * M + a];
                }
            }
    }
    for (int k = 0; k < C; ++k) {
        const int64_t i = int64_t(r.t.local + k) * kBT + threadIdx.x;
        if (k < r.c && i < p.N)
            for (int a = 0; a < M; ++a) {
                float q = rv[k][a];
                if (!first) {
                    q = pv[k][a];
                    q += -om * vv[k][a];
                    q *= bt;
                    q += rv[k][a];
                }
                p.v32[2][i * M + a] = q;
            }
    }
}

template <int M, int C>
__device__ __forceinline__ void run_c32(const BatchProblem& p, const TileRun& r, float* lds, ReduceWork w)
{
    const float al = float(p.ks->alpha);
    float acc[C][1];
    float xv[C][M], pv[C][M], rv[C][M], vv[C][M];
    const float* dir = p.nlev > 1 ? p.v32[6] : p.v32[2];   // x moves along M^-1 p with multigrid
    for (int k = 0; k < C; ++k) {
        const int64_t i = int64_t(r.t.local + k) * kBT + threadIdx.x;
        if (k < r.c && i < p.N)
            for (int a = 0; a < M; ++a) {
                const int64_t q = i * M + a;
                xv[k][a] = p.v32[5][q];
                pv[k][a] = dir[q];
                rv[k][a] = p.v32[0][q];
                vv[k][a] = p.v32[3][q];
            }
    }
    for (int k = 0; k < C; ++k) {
        acc[k][0] = 0.f;
        const int64_t i = int64_t(r.t.local + k) * kBT + threadIdx.x;
        if (k < r.c && i < p.N)
            for (int a = 0; a < M; ++a) {
                const int64_t q = i * M + a;
                float xn = xv[k][a];
                xn += al * pv[k][a];
                p.v32[5][q] = xn;
                const float ri = rv[k][a] + (-al) * vv[k][a];
                p.v32[0][q] = ri;
                acc[k][0] += ri * ri;
            }
    }
    run_partials<1, C>(acc, r.c, lds, w, r.tix);
}

template <int M, int C>
__device__ __forceinline__ void run_e32(const BatchProblem& p, const TileRun& r, float* lds, ReduceWork w)
{
    const float om = float(p.ks->omega);
    float acc[C][2];
    float xv[C][M], sv[C][M], tv[C][M], rt[C][M], dv[C][M];
    const bool mg = p.nlev > 1;   // x moves along M^-1 s with multigrid
    for (int k = 0; k < C; ++k) {
        const int64_t i = int64_t(r.t.local + k) * kBT + threadIdx.x;
        if (k < r.c && i < p.N)
            for (int a = 0; a < M; ++a) {
                const int64_t q = i * M + a;
                xv[k][a] = p.v32[5][q];
                sv[k][a] = p.v32[0][q];
                tv[k][a] = p.v32[4][q];
                rt[k][a] = p.v32[1][q];
                dv[k][a] = mg ? p.v32[6][q] : sv[k][a];
            }
    }
    for (int k = 0; k < C; ++k) {
        acc[k][0] = acc[k][1] = 0.f;
        const int64_t i = int64_t(r.t.local + k) * kBT + threadIdx.x;
        if (k < r.c && i < p.N)
            for (int a = 0; a < M; ++a) {
                const int64_t q = i * M + a;
                float xn = xv[k][a];
                xn += om * dv[k][a];
                p.v32[5][q] = xn;
                const float ri = sv[k][a] + (-om) * tv[k][a];
                p.v32[0][q] = ri;
                acc[k][0] += ri * ri;
                acc[k][1] += rt[k][a] * ri;
            }
    }
    run_partials<2, C>(acc, r.c, lds, w, r.tix);
}

// the FP64 form (fallback problems) runs tile by tile through the tile_* functions
#define DMX_KRYLOV_RUNS(FP32_CALL, FP64_CALL)                                                        \
    int64_t u, u1;                                                                                \
    krylov_range(W, u, u1);                                                                       \
    TileRun r;                                                                                    \
    while (next_run<C>(W, u, u1, probs, r)) {                                                        \
        const BatchProblem& p = probs[r.t.prob];                                                  \
        if (!krylov_active(st[r.t.prob], p.ks)) continue;                                         \
        if (st[r.t.prob].prec64) {                                                                \
            for (int k = 0; k < r.c; ++k) {                                                       \
                const TileRef t{r.t.prob, r.t.local + k};                                         \
                const int64_t tix = r.tix + k;                                                    \
                FP64_CALL;                                                                        \
            }                                                                                     \
        } else {                                                                                  \
            FP32_CALL;                                                                            \
        }                                                                                         \
    }

template <int M, int C>
__global__ void __launch_bounds__(kBT) k_t_a(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W)
{
    DMX_KRYLOV_RUNS((run_a32<M, C>(p, r)), (tile_a<M>(probs, st, t, tix)))
}

// ---- multigrid preconditioner (mg.hh) --------------------------------------------------------
// The level kernels loop over static tables of the boxes of their level (the problems that have
// the level). A problem takes part while its FP32 BiCGSTAB runs; the coarse operators are built
// when its operator was just assembled. Levels below the problem's tail level run box-tiled
// (a thread per fine cell, the children of an aggregate summed in LDS); the tail levels run in
// one workgroup per problem, so a V-cycle costs 2 * mgTail + 1 launches.
static_assert(kBT >= kMgBoxCells, "multigrid launch shape");

__device__ __forceinline__ bool mg_active(const ProbState& s, const BatchProblem& p)
{
    return p.nlev > 1 && !s.prec64 && krylov_active(s, p.ks);
}

// Galerkin operator of level lvl >= 1: a thread per coarse cell of the level's boxes
template <int M>
__global__ void __launch_bounds__(kBT) k_mg_galerkin(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st,
                                                     const TileRef* __restrict__ tiles, int64_t ntiles, int lvl)
{
    for (int64_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const TileRef t = tiles[k];
        const BatchProblem& p = probs[t.prob];
        if (!in_asm(st[t.prob].phase)) continue;
        const MgLevel& c = p.lev[lvl];
        int cidx[3];
        if (!mg_box_cell(c.g, c.box, t.local, threadIdx.x, cidx)) continue;
        const int64_t C = mg_index(c.g, cidx[0], cidx[1], cidx[2]);
        if (lvl == 1) mg_galerkin_jacobian<M>(p.g, p.rowptr, p.J, p.lev[0].Dr, c, cidx, C);
        else mg_galerkin<M>(p.lev[lvl - 1], c, cidx, C);
    }
}

// V-cycle input of level lvl (0: p or s by `which`) and its output (level 0: y)
__device__ __forceinline__ const float* mg_rhs(const BatchProblem& p, int lvl, int which)
{
    return lvl == 0 ? (which == 0 ? p.v32[2] : p.v32[0]) : p.lev[lvl].f;
}

__device__ __forceinline__ float* mg_out(const BatchProblem& p, int lvl)
{
    return lvl == 0 ? p.v32[6] : p.lev[lvl].f;
}

// level lvl of a problem with the pointers that level 0 (L0: unit diagonal, bfloat16 operator,
// restriction weights) or a coarse level (explicit D, D^-1, FP32 operator) does not use set
// to null at compile time, so the sweeps' branches on them fold away
template <bool L0>
__device__ __forceinline__ MgLevel mg_level(const BatchProblem& p, int lvl)
{
    MgLevel L = p.lev[lvl];
    if constexpr (L0) {
        L.D = nullptr;
        L.Dinv = nullptr;
    } else {
        L.Dr = nullptr;
        L.S16 = nullptr;
    }
    return L;
}

// down-sweep from level lvl to lvl+1 over the boxes of level lvl (L0: lvl == 0)
template <int M, bool L0>
__global__ void __launch_bounds__(kBT) k_mg_down(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st,
                                                 const TileRef* __restrict__ tiles, int64_t ntiles, int lvl, int which)
{
    for (int64_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const TileRef t = tiles[k];
        const BatchProblem& p = probs[t.prob];
        if (!mg_active(st[t.prob], p) || lvl >= p.mgTail) continue;   // uniform in the workgroup
        mg_down_box<M>(mg_level<L0>(p, lvl), p.lev[lvl + 1], mg_rhs(p, lvl, which), t.local);
    }
}

// up-sweep at level lvl over its boxes (a thread per fine cell; L0: lvl == 0)
template <int M, bool L0>
__global__ void __launch_bounds__(kBT) k_mg_up(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st,
                                               const TileRef* __restrict__ tiles, int64_t ntiles, int lvl, int which)
{
    for (int64_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const TileRef t = tiles[k];
        const BatchProblem& p = probs[t.prob];
        if (!mg_active(st[t.prob], p) || lvl >= p.mgTail) continue;
        mg_up_box<M>(mg_level<L0>(p, lvl), p.lev[lvl + 1].g, p.lev[lvl + 1].f, mg_rhs(p, lvl, which), mg_out(p, lvl),
                     t.local);
    }
}

// The tail of the V-cycle, one problem per workgroup (mg_tail_cycle from the problem's tail level)
template <int M>
__global__ void __launch_bounds__(kMgTailThreads) k_mg_tail(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st,
                                                            const int32_t* __restrict__ plist, int which)
{
    __shared__ float lev_lds[kMgTailLdsCells * mg_tail_row<M>()];
    __shared__ float r_lds[kMgTailCells * M];
    const int32_t pi = plist[blockIdx.x];
    const BatchProblem& p = probs[pi];
    if (!mg_active(st[pi], p)) return;   // uniform in the workgroup
    mg_tail_cycle<M>(p.lev, p.nlev, p.mgTail, mg_rhs(p, p.mgTail, which), mg_out(p, p.mgTail), lev_lds, r_lds);
}

// one tile per workgroup and pass (the stencil gathers of several tiles do not overlap)
template <int M>
__global__ void __launch_bounds__(kBT) k_t_spmv(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W, ReduceWork w, int which)
{
    __shared__ double lds4d[4];
    __shared__ float lds4f[4];
    int64_t u, u1;
    krylov_range(W, u, u1);
    for (; u < u1; ++u) {
        const TileRef t = W.kmap[u];
        tile_spmv<M>(probs, st, w, which, t, int64_t(probs[t.prob].tile0) + t.local, lds4d, lds4f);
    }
}

template <int M, int C>
__global__ void __launch_bounds__(kBT) k_t_c(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W, ReduceWork w)
{
    __shared__ double lds4d[4];
    __shared__ float lds4f[4];
    __shared__ float ldsr[C * 4];
    DMX_KRYLOV_RUNS((run_c32<M, C>(p, r, ldsr, w)), (tile_c<M>(probs, st, w, t, tix, lds4d, lds4f)))
}

template <int M, int C>
__global__ void __launch_bounds__(kBT) k_t_e(const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st, Work W, ReduceWork w)
{
    __shared__ double lds4d[4];
    __shared__ float lds4f[4];
    __shared__ float ldsr[C * 2 * 4];
    DMX_KRYLOV_RUNS((run_e32<M, C>(p, r, ldsr, w)), (tile_e<M>(probs, st, w, t, tix, lds4d, lds4f)))
}
#undef DMX_KRYLOV_RUNS

// ---- per-problem reductions and control (one wave per problem of a kind) ----------------------
// sum (or max) of a problem's tile partials in tile order; every lane returns the total
constexpr int kCtlThreads = 1024;

// fixed-order reduction over the 16 waves of a 1024-thread workgroup
__device__ __forceinline__ double wg16_reduce(double v, bool max, double* lds16)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double y = __shfl_xor(v, o, 64);
        v = max ? __builtin_fmax(v, y) : v + y;
    }
    if ((threadIdx.x & 63) == 0) lds16[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = lds16[0];
    for (int k = 1; k < 16; ++k) r = max ? __builtin_fmax(r, lds16[k]) : r + lds16[k];
    __syncthreads();
    return r;
}

// sum (or max) of a problem's tile partials: each thread takes 4 independent strided streams
// (the loads overlap), combined in a fixed order, then the workgroup reduction
__device__ __forceinline__ double problem_total(const BatchProblem& p, ReduceWork w, int k, bool max, double* lds16)
{
    const double* src = w.partial + size_t(k) * w.cap + p.tile0;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    int32_t b = threadIdx.x;
    for (; b + 3 * kCtlThreads < p.ntiles; b += 4 * kCtlThreads)
        for (int q = 0; q < 4; ++q) {
            const double v = src[b + q * kCtlThreads];
            a[q] = max ? __builtin_fmax(a[q], v) : a[q] + v;
        }
    for (int q = 0; b < p.ntiles; b += kCtlThreads, ++q) {
        const double v = src[b];
        a[q] = max ? __builtin_fmax(a[q], v) : a[q] + v;
    }
    const double acc = max ? __builtin_fmax(__builtin_fmax(a[0], a[1]), __builtin_fmax(a[2], a[3]))
                           : (a[0] + a[1]) + (a[2] + a[3]);
    return wg16_reduce(acc, max, lds16);
}

// BiCGSTAB scalar steps of the problems in PH_KRYLOV
__global__ void __launch_bounds__(kCtlThreads) k_t_fin(int op, const BatchProblem* __restrict__ probs, const ProbState* __restrict__ st,
                                                       const int32_t* __restrict__ plist, ReduceWork w, int32_t maxit)
{
    __shared__ double lds4[16];
    const int32_t pi = plist[blockIdx.x];
    const BatchProblem& p = probs[pi];
    if (!krylov_active(st[pi], p.ks)) return;
    const int nv = (op == FIN_OMEGA || op == FIN_HALF2) ? 2 : 1;
    double tot[2] = {0.0, 0.0};
    for (int k = 0; k < nv; ++k) tot[k] = problem_total(p, w, k, false, lds4);
    if (threadIdx.x != 0) return;
    fin_apply(op, tot, p.ks, nullptr, 0.0, maxit);
}

// Newton failure: NumericalProblem caught by solve(vars, timeLoop) (newtonsolver.hh:309-355):
// retry the time step with dt * retryFactor from the previous time level, or give up
__device__ void newton_fail(ProbState& s, const BatchParams& prm)
{
    s.wastedSteps += s.steps;
    s.steps = 0;
    if (!s.linearPde && s.retries < prm.maxDiv) {
        s.retries += 1;
        s.dtRetries += 1;
        s.dt *= prm.retryFactor;
        s.phase = PH_RESET;
    } else {
        s.status = -1;
        s.phase = PH_IDLE;
    }
}

// the refinement loop asks for the next FP32 solve of ir (scaled to unit norm)
__device__ __forceinline__ void next_solve(ProbState& s)
{
    s.scale = 1.0 / s.res;
    s.corr = s.res;
    s.phase = PH_INIT;
}

enum CtlStage : int32_t { CTL_NEWTON = 0, CTL_ASM = 1, CTL_INIT = 2, CTL_KRYLOV = 3, CTL_CORRECT = 4 };

__global__ void __launch_bounds__(kCtlThreads) k_t_ctl(int stage, const BatchProblem* __restrict__ probs, ProbState* __restrict__ st,
                                                       const int32_t* __restrict__ plist, ReduceWork w, BatchParams prm)
{
    __shared__ double lds4[16];
    const int32_t pi = plist[blockIdx.x];
    const BatchProblem& p = probs[pi];
    ProbState* sp = st + pi;
    const int32_t ph = sp->phase;
    switch (stage) {
    case CTL_NEWTON: {
        if (ph != PH_UPDATE) return;
        const double shift = problem_total(p, w, 0, true, lds4);
        if (threadIdx.x != 0) return;
        ProbState s = *sp;
        s.shift = shift;
        s.steps += 1;
        const bool converged = shift <= prm.maxRelShift;
        bool success = false, failure = false;
        if (s.linearPde) {
            success = true;
        } else if (s.steps < prm.minSteps) {
        } else if (converged) {
            success = true;
        } else if (s.steps >= prm.maxSteps) {
            failure = !(s.shift * 4.0 < s.lastShift);
        }
        if (success) {
            const int32_t used = s.steps;
            s.newtonSteps += s.steps;
            s.steps = 0;
            s.retries = 0;
            s.t += s.dt;
            s.timeSteps += 1;
            if (s.linearPde || s.t >= s.tEnd) {
                s.status = 1;
                s.phase = PH_IDLE;
            } else {
                // fixed dt, or the suggested one capped by TimeLoop::maxTimeStepSize
                // (common/timeloop.hh:320-324, 404-411)
                if (prm.adaptive) {
                    const double cap = __builtin_fmin(prm.maxDt > 0.0 ? prm.maxDt : s.tEnd, s.tEnd - s.t);
                    s.dt = __builtin_fmin(suggest_time_step(s.dt, used, prm.targetSteps), cap);
                } else {
                    s.dt = __builtin_fmin(s.dtNom, s.tEnd - s.t);
                }
                s.phase = PH_ADVANCE;
            }
        } else if (failure) {
            newton_fail(s, prm);
        } else {
            s.phase = PH_ASM;
        }
        *sp = s;
        return;
    }
    case CTL_ASM: {
        if (!in_asm(ph) && ph != PH_RESTART) return;
        const double nr2 = problem_total(p, w, 0, false, lds4);
        if (threadIdx.x != 0) return;
        ProbState s = *sp;
        if (ph != PH_RESTART) s.lastShift = s.shift;  // newtonBeginStep (newtonsolver.hh:451-453)
        s.normb = __builtin_sqrt(nr2);
        s.res = s.normb;
        s.resPrev = s.normb;
        s.ref = 0;
        if (!__builtin_isfinite(s.normb)) newton_fail(s, prm);
        else if (s.normb == 0.0) s.phase = PH_UPDATE;  // du = 0 solves the system
        else next_solve(s);
        *sp = s;
        return;
    }
    case CTL_INIT: {
        if (ph != PH_INIT) return;
        double tot[2] = {problem_total(p, w, 0, false, lds4), 0.0};
        if (threadIdx.x != 0) return;
        fin_apply(FIN_INIT, tot, p.ks, nullptr, prm.innerReduction, prm.maxIter);
        sp->phase = PH_KRYLOV;
        return;
    }
    case CTL_KRYLOV: {
        if (ph != PH_KRYLOV || threadIdx.x != 0) return;
        const KrylovScalars* ks = p.ks;
        if (!ks->done) return;
        ProbState s = *sp;
        s.linIts += (ks->half_steps + 1) / 2;
        s.refinements += 1;
        if (ks->status != 0 && ks->half_steps == 0) {  // breakdown before any progress
            if (s.prec64) {
                newton_fail(s, prm);
            } else {
                s.prec64 = 1;
                next_solve(s);
            }
        } else {
            s.phase = PH_CORRECT;
        }
        *sp = s;
        return;
    }
    default: {  // CTL_CORRECT
        if (ph != PH_CORRECT) return;
        const double nr2 = problem_total(p, w, 0, false, lds4);
        if (threadIdx.x != 0) return;
        ProbState s = *sp;
        const double res = __builtin_sqrt(nr2);
        if (!__builtin_isfinite(res)) {
            // an FP32 correction overflowed: restart this linear solve in FP64
            if (s.prec64) newton_fail(s, prm);
            else { s.prec64 = 1; s.phase = PH_RESTART; }
        } else if (res <= s.reduction * s.normb) {
            s.res = res;
            s.phase = PH_UPDATE;
        } else if (s.ref + 1 < prm.maxRef) {
            // the FP32 solves stall when cond(A) * eps32 is not small: two refinements in a row
            // that reduce the residual less than 4x (a healthy one gains ~1/inner_reduction)
            // switch the problem to FP64 solves; one weak refinement (an inner breakdown) does not
            if (!s.prec64) {
                s.stalls = res > 0.25 * s.res ? s.stalls + 1 : 0;
                if (s.stalls >= 2) s.prec64 = 1;
            }
            s.resPrev = s.res;
            s.res = res;
            s.ref += 1;
            next_solve(s);
        } else {
            newton_fail(s, prm);  // the linear solver did not converge
        }
        *sp = s;
        return;
    }
    }
}

} // namespace

} // namespace dmx

// =============================================================================================
// host executor
// =============================================================================================
using dmx::BatchParams;
using dmx::BatchProblem;
using dmx::ProbState;
using dmx::TileRef;

struct dmx_batch {
    struct Kind {
        int dim = 0, m = 0;
        int64_t ntiles = 0;               // tiles of all problems of the kind
        std::vector<int32_t> probs;       // problem indices
        bool numeric = false, analytic = false;  // members with forward-difference / hand-coded Jacobians
        int32_t* d_plist = nullptr;
        int32_t* d_work = nullptr;        // work lists (dmx::Work)
        TileRef* d_kmap = nullptr;        // BiCGSTAB tile map
        // multigrid: static tables of the boxes of every level (problems with that level),
        // level l at [mgOff[l], mgOff[l+1]); levels below mgTiled run box-tiled
        int mgLevels = 1;
        int mgTiled = 0;
        std::vector<int64_t> mgOff;
        TileRef* d_mgTiles = nullptr;
    };
    struct Member {
        dmx_context* ctx = nullptr;
        int32_t slot = -1;                // index in the device arrays of the current layout, -1: new
        ProbState init{};                 // initial state of a new member
    };

    int device = 0;
    hipStream_t stream = nullptr;
    BatchParams prm{};
    double reductionDefault = 1e-10;
    std::vector<Member> mem;
    std::vector<Kind> kinds;
    std::vector<BatchProblem> h_probs;
    int32_t nslots = 0;                   // members in the device arrays
    BatchProblem* d_probs = nullptr;
    ProbState* d_st = nullptr;
    ProbState* h_st = nullptr;            // pinned mirror, refreshed after every step
    KrylovScalars* d_ks = nullptr;
    KrylovScalars* h_ks = nullptr;        // pinned
    int32_t* d_fail = nullptr;
    dmx::ReduceWork rw{};
    int64_t capProbs = 0;
    bool dirty = true;
    bool useGraph = true;
    // launch shape of the BiCGSTAB kernels (dmx_batch_autotune): tiles per pass of the vector
    // kernels, their grid, and the SpMV grid
    struct Tune {
        int coarse = 4;
        int64_t vecGrid = 2048;
        int64_t spmvGrid = 8192;
    } tune;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string err;
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
    if (k.d_mgTiles) (void)hipFree(k.d_mgTiles);
    k.d_mgTiles = nullptr;
    if (k.d_work) (void)hipFree(k.d_work);
    if (k.d_kmap) (void)hipFree(k.d_kmap);
    k.d_kmap = nullptr;
    if (k.d_plist) (void)hipFree(k.d_plist);
    k.d_work = nullptr;
    k.d_plist = nullptr;
}

void drop_graph(dmx_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// make sure the context has the arrays of the mixed-precision block-Jacobi solve
int prepare_context(dmx_batch* b, dmx_context* c)
{
    const int mm = c->m * c->m;
    auto need = [&](auto** p, int64_t cnt) -> int {
        if (*p) return DMX_OK;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(p), sizeof(**p) * size_t(std::max<int64_t>(cnt, 1)));
        if (e != hipSuccess) { b->err = std::string("hipMalloc: ") + hipGetErrorString(e); return DMX_ERR_NOMEM; }
        c->devBytes += int64_t(sizeof(**p)) * std::max<int64_t>(cnt, 1);
        return DMX_OK;
    };
    B_TRY(need(&c->S32, int64_t(2) * c->g.dim * mm * c->N));
    B_TRY(need(&c->Dinv32, c->N * mm));
    B_TRY(need(&c->Dinv64, c->N * mm));
    B_TRY(need(&c->ir_r, c->n));
    B_TRY(need(&c->kvb, 7 * c->n));
    if (b->prm.mg) B_TRY(need(&c->mgBuf, mg_floats(c->g, c->m)));
    return DMX_OK;
}

BatchProblem describe(dmx_batch* b, dmx_context* c, int pi)
{
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
    P.vv_d = c->vvd;
    P.mat = c->mat;
    P.analytic = c->desc.diff_method == DMX_DIFF_ANALYTIC ? 1 : 0;
    P.rowptr = c->rowptr;
    P.r = c->r;
    P.J = c->J;
    P.S32 = c->S32;
    P.Dinv32 = c->Dinv32;
    P.Dinv64 = c->Dinv64;
    P.fail = b->d_fail;
    P.dt = c->dt;
    P.rowBase = 0;
    p.vvp = c->vvp;
    p.vvu = c->vvu;
    p.vvd = c->vvd;
    p.rowptr = c->rowptr;
    p.colidx = c->colidx;
    p.S = c->S32;
    p.Dinv32 = c->Dinv32;
    p.J = c->J;
    for (int i = 0; i < 8; ++i) p.v32[i] = reinterpret_cast<float*>(c->kvb) + i * c->n;
    for (int i = 0; i < 7; ++i) p.v64[i] = c->kvb + i * c->n;
    p.Dinv64 = c->Dinv64;
    p.r = c->r;
    p.du = c->du;
    p.ir = c->ir_r;
    p.u = c->u;
    p.u_last = c->u_last;
    p.uprev = c->uprev;
    p.ks = b->d_ks + pi;
    p.N = c->N;
    p.n = c->n;
    p.nlev = 1;
    p.mgTail = 0;
    p.lev[0] = MgLevel{c->g, mg_boxes(c->g), c->S32, nullptr, nullptr, nullptr, p.v32[7], nullptr, nullptr};
    if (b->prm.mg && mg_num_levels(c->g) > 1) {
        // the context's multigrid buffer: level 0's restriction weights, then S, D, Dinv, x, f
        // of every coarse level
        const int nl = mg_num_levels(c->g);
        float* q = c->mgBuf;
        const int mm = c->m * c->m;
        p.lev[0].S16 = reinterpret_cast<uint16_t*>(q);
        q += int64_t(c->g.dim) * mm * c->g.N;
        p.lev[0].Dr = q;
        q += mm * c->g.N;
        for (int l = 1; l < nl; ++l) {
            const GridDesc gc = mg_coarse_grid(p.lev[l - 1].g);
            MgLevel& L = p.lev[l];
            L.g = gc;
            L.box = mg_boxes(gc);
            L.S = q; q += int64_t(2) * gc.dim * mm * gc.N;
            L.D = q; q += mm * gc.N;
            L.Dinv = q; q += mm * gc.N;
            L.Dr = nullptr;
            L.S16 = nullptr;
            L.x = q; q += c->m * gc.N;
            L.f = q; q += c->m * gc.N;
        }
        p.mgTail = mg_tail_level(c->g, nl);
        p.nlev = nl;
    }
    return p;
}

// tile table of a kind, and its grid padded to a multiple of the XCD count
// work lists of a kind, and the grid of its phase kernels: a multiple of the XCD count, at most
// kMaxGrid workgroups (each loops over its share of the active tiles); the BiCGSTAB kernels'
// grids come from the batch's launch shape (dmx_batch::Tune)
constexpr int64_t kMaxGrid = 2048;
dmx::Work work_of(const dmx_batch::Kind& K)
{
    const int32_t cap = int32_t(K.probs.size());
    return dmx::Work{K.d_work, K.d_work + dmx::kGroups, K.d_work + 2 * dmx::kGroups,
                     K.d_work + 2 * dmx::kGroups + int64_t(dmx::kGroups) * cap, cap, K.d_kmap};
}
unsigned xcd_grid(const dmx_batch::Kind& K, int64_t maxGrid = kMaxGrid)
{
    const int64_t nb = std::min<int64_t>(K.ntiles, maxGrid);
    return unsigned((nb + dmx::kXcds - 1) / dmx::kXcds * dmx::kXcds);
}

template <class F>
int each_kind(dmx_batch* b, F&& f)
{
    for (auto& K : b->kinds)
        if (!K.probs.empty()) {
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
        const dim3 g_(xcd_grid(K)), t_(dmx::kBT);                                                     \
        if (K.m == 1) kernel<1><<<g_, t_, 0, b->stream>>>(__VA_ARGS__);                               \
        else kernel<2><<<g_, t_, 0, b->stream>>>(__VA_ARGS__);                                        \
        return hipGetLastError();                                                                     \
    }()

// vector kernels: instantiation by unknowns per cell and tiles per pass
#define KLAUNCH_VEC(kernel, K, ...)                                                                   \
    [&]() -> hipError_t {                                                                             \
        const dim3 g_(xcd_grid(K, b->tune.vecGrid)), t_(dmx::kBT);                                   \
        switch (K.m * 16 + b->tune.coarse) {                                                          \
        case 16 + 1: kernel<1, 1><<<g_, t_, 0, b->stream>>>(__VA_ARGS__); break;                     \
        case 16 + 2: kernel<1, 2><<<g_, t_, 0, b->stream>>>(__VA_ARGS__); break;                     \
        case 16 + 4: kernel<1, 4><<<g_, t_, 0, b->stream>>>(__VA_ARGS__); break;                     \
        case 16 + 8: kernel<1, 8><<<g_, t_, 0, b->stream>>>(__VA_ARGS__); break;                     \
        case 32 + 1: kernel<2, 1><<<g_, t_, 0, b->stream>>>(__VA_ARGS__); break;                     \
        case 32 + 2: kernel<2, 2><<<g_, t_, 0, b->stream>>>(__VA_ARGS__); break;                     \
        case 32 + 4: kernel<2, 4><<<g_, t_, 0, b->stream>>>(__VA_ARGS__); break;                     \
        default: kernel<2, 8><<<g_, t_, 0, b->stream>>>(__VA_ARGS__); break;                         \
        }                                                                                             \
        return hipGetLastError();                                                                     \
    }()

#define KLAUNCH_SPMV(kernel, K, ...)                                                                  \
    [&]() -> hipError_t {                                                                             \
        const dim3 g_(xcd_grid(K, b->tune.spmvGrid)), t_(dmx::kBT);                                  \
        if (K.m == 1) kernel<1><<<g_, t_, 0, b->stream>>>(__VA_ARGS__);                               \
        else kernel<2><<<g_, t_, 0, b->stream>>>(__VA_ARGS__);                                        \
        return hipGetLastError();                                                                     \
    }()

hipError_t plan(dmx_batch* b, dmx_batch::Kind& K)
{
    hipLaunchKernelGGL(dmx::k_t_plan, dim3(1), dim3(1024), 0, b->stream, b->d_probs, b->d_st, K.d_plist, work_of(K));
    return hipGetLastError();
}

// plan, then the BiCGSTAB group's tile map
hipError_t plan_krylov(dmx_batch* b, dmx_batch::Kind& K)
{
    hipError_t e = plan(b, K);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dmx::k_t_expand, dim3(unsigned(K.probs.size())), dim3(256), 0, b->stream, b->d_probs, work_of(K));
    return hipGetLastError();
}

hipError_t ctl(dmx_batch* b, dmx_batch::Kind& K, int stage)
{
    hipLaunchKernelGGL(dmx::k_t_ctl, dim3(unsigned(K.probs.size())), dim3(dmx::kCtlThreads), 0, b->stream, stage, b->d_probs, b->d_st,
                       K.d_plist, b->rw, b->prm);
    return hipGetLastError();
}

hipError_t fin(dmx_batch* b, dmx_batch::Kind& K, int op)
{
    hipLaunchKernelGGL(dmx::k_t_fin, dim3(unsigned(K.probs.size())), dim3(dmx::kCtlThreads), 0, b->stream, op, b->d_probs, b->d_st,
                       K.d_plist, b->rw, b->prm.maxIter);
    return hipGetLastError();
}

int launch_assembly(dmx_batch* b)
{
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_t_state, K, b->d_probs, b->d_st, work_of(K)); }));
    return each_kind(b, [&](auto& K) -> hipError_t {
        const dim3 g(xcd_grid(K)), t(dmx::kBT);
#define DMX_T_ASM(D, MM, AN) \
    hipLaunchKernelGGL((dmx::k_t_assemble<D, MM, AN>), g, t, 0, b->stream, b->d_probs, b->d_st, work_of(K))
#define DMX_T_ASM_KIND(D, MM)                   \
    if (K.dim == D && K.m == MM) {              \
        if (K.numeric) DMX_T_ASM(D, MM, false); \
        if (K.analytic) DMX_T_ASM(D, MM, true); \
    }
        DMX_T_ASM_KIND(2, 1)
        DMX_T_ASM_KIND(2, 2)
        DMX_T_ASM_KIND(3, 1)
        DMX_T_ASM_KIND(3, 2)
#undef DMX_T_ASM_KIND
#undef DMX_T_ASM
        return hipGetLastError();
    });
}

// multigrid: Galerkin coarse operators of freshly assembled Jacobians
int launch_mg_setup(dmx_batch* b)
{
    return each_kind(b, [&](auto& K) -> hipError_t {
        for (int l = 1; l < K.mgLevels; ++l) {
            const int64_t t0 = K.mgOff[l], nt = K.mgOff[l + 1] - K.mgOff[l];
            if (nt <= 0) continue;
            const dim3 g(unsigned(std::min<int64_t>(nt, kMaxGrid))), t(dmx::kBT);
            if (K.m == 1) dmx::k_mg_galerkin<1><<<g, t, 0, b->stream>>>(b->d_probs, b->d_st, K.d_mgTiles + t0, nt, l);
            else dmx::k_mg_galerkin<2><<<g, t, 0, b->stream>>>(b->d_probs, b->d_st, K.d_mgTiles + t0, nt, l);
        }
        return hipGetLastError();
    });
}

// one box-tiled down- (up = false) or up-sweep of level l of a kind
template <class KindT>
hipError_t launch_mg_sweep(dmx_batch* b, KindT& K, int l, int which, bool up)
{
    const int64_t t0 = K.mgOff[l], nt = K.mgOff[l + 1] - K.mgOff[l];
    if (nt <= 0) return hipSuccess;
    const dim3 g(unsigned(std::min<int64_t>(nt, kMaxGrid))), t(dmx::kBT);
    const TileRef* tl = K.d_mgTiles + t0;
    auto go = [&](auto kernel) { kernel<<<g, t, 0, b->stream>>>(b->d_probs, b->d_st, tl, nt, l, which); };
    const bool l0 = l == 0;
    if (K.m == 1) {
        if (up) l0 ? go(dmx::k_mg_up<1, true>) : go(dmx::k_mg_up<1, false>);
        else l0 ? go(dmx::k_mg_down<1, true>) : go(dmx::k_mg_down<1, false>);
    } else {
        if (up) l0 ? go(dmx::k_mg_up<2, true>) : go(dmx::k_mg_up<2, false>);
        else l0 ? go(dmx::k_mg_down<2, true>) : go(dmx::k_mg_down<2, false>);
    }
    return hipGetLastError();
}

// y = M^-1 (p or s): box-tiled down-sweeps, the tails, box-tiled up-sweeps
int launch_mg_vcycle(dmx_batch* b, int which)
{
    return each_kind(b, [&](auto& K) -> hipError_t {
        if (K.mgLevels < 2) return hipSuccess;
        for (int l = 0; l < K.mgTiled; ++l)
            if (const hipError_t e = launch_mg_sweep(b, K, l, which, false); e != hipSuccess) return e;
        const dim3 gt(unsigned(K.probs.size())), tt(dmx::kMgTailThreads);
        if (K.m == 1) dmx::k_mg_tail<1><<<gt, tt, 0, b->stream>>>(b->d_probs, b->d_st, K.d_plist, which);
        else dmx::k_mg_tail<2><<<gt, tt, 0, b->stream>>>(b->d_probs, b->d_st, K.d_plist, which);
        for (int l = K.mgTiled - 1; l >= 0; --l)
            if (const hipError_t e = launch_mg_sweep(b, K, l, which, true); e != hipSuccess) return e;
        return hipGetLastError();
    });
}

int launch_bicgstab_iteration(dmx_batch* b)
{
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH_VEC(dmx::k_t_a, K, b->d_probs, b->d_st, work_of(K)); }));
    if (b->prm.mg) B_TRY(launch_mg_vcycle(b, 0));
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH_SPMV(dmx::k_t_spmv, K, b->d_probs, b->d_st, work_of(K), b->rw, 0); }));
    B_TRY(each_kind(b, [&](auto& K) { return fin(b, K, dmx::FIN_H); }));
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH_VEC(dmx::k_t_c, K, b->d_probs, b->d_st, work_of(K), b->rw); }));
    B_TRY(each_kind(b, [&](auto& K) { return fin(b, K, dmx::FIN_HALF1); }));
    if (b->prm.mg) B_TRY(launch_mg_vcycle(b, 1));
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH_SPMV(dmx::k_t_spmv, K, b->d_probs, b->d_st, work_of(K), b->rw, 1); }));
    B_TRY(each_kind(b, [&](auto& K) { return fin(b, K, dmx::FIN_OMEGA); }));
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH_VEC(dmx::k_t_e, K, b->d_probs, b->d_st, work_of(K), b->rw); }));
    return each_kind(b, [&](auto& K) { return fin(b, K, dmx::FIN_HALF2); });
}

// the launches of one tick (independent of the problems' states: captured as a graph)
int enqueue_tick(dmx_batch* b)
{
    using namespace dmx;
    auto planAll = [&]() { return each_kind(b, [&](auto& K) { return plan(b, K); }); };
    B_TRY(planAll());
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(k_t_update, K, b->d_probs, b->d_st, work_of(K), b->rw); }));
    B_TRY(each_kind(b, [&](auto& K) { return ctl(b, K, CTL_NEWTON); }));
    B_TRY(planAll());
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(k_t_level, K, b->d_probs, b->d_st, work_of(K)); }));
    B_TRY(launch_assembly(b));
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(k_t_begin, K, b->d_probs, b->d_st, work_of(K), b->rw); }));
    if (b->prm.mg) B_TRY(launch_mg_setup(b));
    B_TRY(each_kind(b, [&](auto& K) { return ctl(b, K, CTL_ASM); }));
    B_TRY(planAll());
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(k_t_init, K, b->d_probs, b->d_st, work_of(K), b->rw); }));
    B_TRY(each_kind(b, [&](auto& K) { return ctl(b, K, CTL_INIT); }));
    B_TRY(each_kind(b, [&](auto& K) { return plan_krylov(b, K); }));
    for (int it = 0; it < kTickIters; ++it) B_TRY(launch_bicgstab_iteration(b));
    B_TRY(each_kind(b, [&](auto& K) { return ctl(b, K, CTL_KRYLOV); }));
    B_TRY(planAll());
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(k_t_correct, K, b->d_probs, b->d_st, work_of(K)); }));
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(k_t_residual, K, b->d_probs, b->d_st, work_of(K), b->rw); }));
    return each_kind(b, [&](auto& K) { return ctl(b, K, CTL_CORRECT); });
}

// (re)capture the tick as a graph with the current members and launch shape
int capture_tick(dmx_batch* b)
{
    drop_graph(b);
    if (b->nslots == 0 || !b->useGraph) return DMX_OK;
    B_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_tick(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != DMX_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc == DMX_OK) b->err = std::string("tick capture: ") + hipGetErrorString(e);
        return rc == DMX_OK ? DMX_ERR_DEVICE : rc;
    }
    b->graph = g;
    B_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return DMX_OK;
}

// lay the members out in the device arrays (kinds, tiles, descriptors), carrying the states
// and Krylov scalars of the members that were already there, and capture the tick
int rebuild(dmx_batch* b)
{
    B_HIP(hipStreamSynchronize(b->stream));
    drop_graph(b);
    const int P = int(b->mem.size());
    // current device states of the members already laid out
    std::vector<ProbState> oldSt(size_t(b->nslots));
    std::vector<KrylovScalars> oldKs(size_t(b->nslots));
    if (b->nslots > 0) {
        B_HIP(hipMemcpy(oldSt.data(), b->d_st, sizeof(ProbState) * b->nslots, hipMemcpyDeviceToHost));
        B_HIP(hipMemcpy(oldKs.data(), b->d_ks, sizeof(KrylovScalars) * b->nslots, hipMemcpyDeviceToHost));
    }
    b->nslots = 0;
    if (P > b->capProbs) {
        for (void* p : {(void*)b->d_probs, (void*)b->d_st, (void*)b->d_ks})
            if (p) (void)hipFree(p);
        if (b->h_st) (void)hipHostFree(b->h_st);
        if (b->h_ks) (void)hipHostFree(b->h_ks);
        b->d_probs = nullptr;
        b->d_st = nullptr;
        b->d_ks = nullptr;
        b->h_st = nullptr;
        b->h_ks = nullptr;
        b->capProbs = std::max<int64_t>(P, 2 * b->capProbs);
        B_TRY(balloc(b, &b->d_probs, b->capProbs));
        B_TRY(balloc(b, &b->d_st, b->capProbs));
        B_TRY(balloc(b, &b->d_ks, b->capProbs));
        B_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_st), sizeof(ProbState) * b->capProbs, hipHostMallocDefault));
        B_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_ks), sizeof(KrylovScalars) * b->capProbs, hipHostMallocDefault));
    }
    for (auto& k : b->kinds) free_kind(k);
    b->kinds.clear();
    b->h_probs.assign(size_t(P), BatchProblem{});
    for (int pi = 0; pi < P; ++pi) {
        auto& m = b->mem[pi];
        dmx_context* c = m.ctx;
        B_TRY(prepare_context(b, c));
        b->h_st[pi] = m.slot >= 0 ? oldSt[size_t(m.slot)] : m.init;
        b->h_ks[pi] = m.slot >= 0 ? oldKs[size_t(m.slot)] : KrylovScalars{};
        m.slot = pi;
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
        BatchProblem p = describe(b, c, pi);
        p.ntiles = int32_t((c->N + dmx::kBT - 1) / dmx::kBT);
        K.ntiles += p.ntiles;
        K.probs.push_back(pi);
        if (p.P.analytic) K.analytic = true;
        else K.numeric = true;
        b->h_probs[pi] = p;
    }
    // partial-sum slots: the tiles of each problem, kind after kind
    int64_t totalTiles = 0;
    for (auto& K : b->kinds) {
        for (int32_t pi : K.probs) {
            b->h_probs[pi].tile0 = int32_t(totalTiles);
            totalTiles += b->h_probs[pi].ntiles;
        }
        const int64_t cap = int64_t(K.probs.size());
        B_TRY(balloc(b, &K.d_plist, cap));
        B_TRY(balloc(b, &K.d_work, 2 * dmx::kGroups + 2 * dmx::kGroups * cap));
        B_TRY(balloc(b, &K.d_kmap, K.ntiles));
        B_HIP(hipMemcpy(K.d_plist, K.probs.data(), sizeof(int32_t) * K.probs.size(), hipMemcpyHostToDevice));
        // multigrid box tables
        K.mgLevels = 1;
        K.mgTiled = 0;
        for (int32_t pi : K.probs) {
            K.mgLevels = std::max(K.mgLevels, int(b->h_probs[pi].nlev));
            if (b->h_probs[pi].nlev > 1) K.mgTiled = std::max(K.mgTiled, int(b->h_probs[pi].mgTail));
        }
        K.mgOff.assign(size_t(K.mgLevels) + 1, 0);
        std::vector<TileRef> tiles;
        for (int l = 0; l < K.mgLevels && K.mgLevels > 1; ++l) {
            K.mgOff[l] = int64_t(tiles.size());
            for (int32_t pi : K.probs) {
                const BatchProblem& bp = b->h_probs[pi];
                if (bp.nlev <= l || bp.nlev < 2) continue;
                for (int32_t t = 0; t < bp.lev[l].box.ntiles; ++t) tiles.push_back(TileRef{pi, t});
            }
        }
        K.mgOff[K.mgLevels] = int64_t(tiles.size());
        if (!tiles.empty()) {
            B_TRY(balloc(b, &K.d_mgTiles, int64_t(tiles.size())));
            B_HIP(hipMemcpy(K.d_mgTiles, tiles.data(), sizeof(TileRef) * tiles.size(), hipMemcpyHostToDevice));
        }
    }
    if (b->rw.cap < totalTiles) {
        if (b->rw.partial) (void)hipFree(b->rw.partial);
        b->rw.partial = nullptr;
        b->rw.cap = totalTiles;
        B_TRY(balloc(b, &b->rw.partial, 2 * b->rw.cap));
    }
    if (P > 0) {
        B_HIP(hipMemcpy(b->d_probs, b->h_probs.data(), sizeof(BatchProblem) * P, hipMemcpyHostToDevice));
        B_HIP(hipMemcpy(b->d_st, b->h_st, sizeof(ProbState) * P, hipMemcpyHostToDevice));
        B_HIP(hipMemcpy(b->d_ks, b->h_ks, sizeof(KrylovScalars) * P, hipMemcpyHostToDevice));
    }
    b->nslots = P;
    b->dirty = false;
    return capture_tick(b);
}

int run_ticks(dmx_batch* b, int ticks, int32_t* nrunning)
{
    if (b->dirty) B_TRY(rebuild(b));
    const int P = b->nslots;
    if (P > 0) {
        for (int k = 0; k < ticks; ++k) {
            if (b->exec) B_HIP(hipGraphLaunch(b->exec, b->stream));
            else B_TRY(enqueue_tick(b));
        }
        B_HIP(hipMemcpyAsync(b->h_st, b->d_st, sizeof(ProbState) * P, hipMemcpyDeviceToHost, b->stream));
        B_HIP(hipStreamSynchronize(b->stream));
    }
    int32_t running = 0;
    for (int pi = 0; pi < P; ++pi) running += b->h_st[pi].status == 0;
    if (nrunning) *nrunning = running;
    return DMX_OK;
}

// every member into a fresh FP32 solve of its last refinement residual that never converges
// (target 0): the state in which the BiCGSTAB kernels are timed; the caller restores the states
int krylov_timing_setup(dmx_batch* b)
{
    const int P = b->nslots;
    for (int pi = 0; pi < P; ++pi) {
        ProbState& s = b->h_st[pi];
        s.phase = dmx::PH_INIT;
        s.prec64 = 0;
        s.scale = s.res > 0.0 ? 1.0 / s.res : 1.0;
    }
    B_HIP(hipMemcpy(b->d_st, b->h_st, sizeof(ProbState) * P, hipMemcpyHostToDevice));
    if (b->prm.mg)   // the preconditioned vector feeds the SpMV timed alone
        for (int pi = 0; pi < P; ++pi)
            B_HIP(hipMemsetAsync(b->h_probs[pi].v32[6], 0, sizeof(float) * size_t(b->h_probs[pi].n), b->stream));
    B_TRY(each_kind(b, [&](auto& K) { return plan(b, K); }));
    BatchParams never = b->prm;
    never.innerReduction = 0.0;
    B_TRY(each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_t_init, K, b->d_probs, b->d_st, work_of(K), b->rw); }));
    B_TRY(each_kind(b, [&](auto& K) -> hipError_t {
        hipLaunchKernelGGL(dmx::k_t_ctl, dim3(unsigned(K.probs.size())), dim3(dmx::kCtlThreads), 0, b->stream, int(dmx::CTL_INIT),
                           b->d_probs, b->d_st, K.d_plist, b->rw, never);
        return hipGetLastError();
    }));
    return each_kind(b, [&](auto& K) { return plan_krylov(b, K); });
}

int find_member(const dmx_batch* b, const dmx_context* c)
{
    for (size_t i = 0; i < b->mem.size(); ++i)
        if (b->mem[i].ctx == c) return int(i);
    return -1;
}

} // namespace

extern "C" {

int dmx_batch_create(int32_t device, const dmx_newton_params* p, dmx_batch** out)
{
    if (!out || !p) return DMX_ERR_INVALID;
    *out = nullptr;
    if (p->linear.max_iter < 1 || p->linear.max_refinements < 1 || !(p->linear.reduction > 0.0) || p->max_steps < 1)
        return DMX_ERR_INVALID;
    // the batch keeps the FP32 operator scaled in place (no previous Jacobian columns to keep)
    if (p->partial_reassembly) return DMX_ERR_UNSUPPORTED;
    auto* b = new dmx_batch();
    b->device = device;
    b->prm.maxRelShift = p->max_rel_shift;
    b->prm.minSteps = p->min_steps;
    b->prm.maxSteps = p->max_steps;
    b->prm.maxDiv = p->max_dt_divisions;
    b->prm.retryFactor = p->retry_factor;
    b->prm.maxIter = p->linear.max_iter;
    b->prm.maxRef = p->linear.max_refinements;
    b->prm.innerReduction = p->linear.inner_reduction;
    b->prm.targetSteps = p->target_steps > 0 ? p->target_steps : 10;
    b->prm.adaptive = p->adaptive_dt;
    b->prm.maxDt = p->max_dt;
    b->prm.mg = p->linear.precond == DMX_PRECOND_AMG ? 1 : 0;
    b->reductionDefault = p->linear.reduction;
    b->useGraph = std::getenv("DMX_NO_GRAPH") == nullptr;
    auto fail = [&](int rc) {
        std::fprintf(stderr, "dmx_batch_create: %s\n", b->err.c_str());
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
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
    drop_graph(b);
    for (auto& k : b->kinds) free_kind(k);
    for (void* p : {(void*)b->d_probs, (void*)b->d_st, (void*)b->d_ks, (void*)b->d_fail, (void*)b->rw.partial})
        if (p) (void)hipFree(p);
    if (b->h_st) (void)hipHostFree(b->h_st);
    if (b->h_ks) (void)hipHostFree(b->h_ks);
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
    if (find_member(b, c) >= 0) {
        b->err = "context already in the batch";
        return DMX_ERR_INVALID;
    }
    B_HIP(hipSetDevice(b->device));
    // the context's own stream may still hold work (uploads)
    B_HIP(hipStreamSynchronize(c->stream));
    dmx_batch::Member m;
    m.ctx = c;
    ProbState& s = m.init;
    s = ProbState{};
    s.linearPde = item->linear_pde != 0;
    if (!s.linearPde) {
        // the time loop starts at the context's solution: uprev <- u
        B_HIP(hipMemcpyAsync(c->uprev, c->u, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
        B_HIP(hipStreamSynchronize(c->stream));
        c->havePrev = true;
    } else {
        c->havePrev = false;
    }
    s.reduction = item->reduction > 0.0 ? item->reduction : b->reductionDefault;
    s.dtNom = item->dt;
    s.tEnd = s.linearPde ? item->dt : item->nsteps * item->dt;
    s.dt = std::min(s.dtNom, s.tEnd);
    if (b->prm.adaptive && b->prm.maxDt > 0.0) s.dt = std::min(s.dt, b->prm.maxDt);
    s.t = 0.0;
    s.phase = s.tEnd > 0.0 ? dmx::PH_ASM : dmx::PH_IDLE;
    s.status = s.tEnd > 0.0 ? 0 : 1;
    c->jacValid = false;  // the batch overwrites the context's Jacobian and vectors
    b->mem.push_back(m);
    b->dirty = true;
    return DMX_OK;
}

int dmx_batch_remove(dmx_batch* b, dmx_context* c)
{
    if (!b || !c) return DMX_ERR_INVALID;
    const int i = find_member(b, c);
    if (i < 0) {
        b->err = "context not in the batch";
        return DMX_ERR_INVALID;
    }
    B_HIP(hipStreamSynchronize(b->stream));
    b->mem.erase(b->mem.begin() + i);
    b->dirty = true;
    return DMX_OK;
}

int dmx_batch_size(const dmx_batch* b) { return b ? int(b->mem.size()) : -1; }

int dmx_batch_step(dmx_batch* b, int32_t ticks, int32_t* nrunning)
{
    if (!b || ticks < 1) return DMX_ERR_INVALID;
    B_HIP(hipSetDevice(b->device));
    return run_ticks(b, ticks, nrunning);
}

int dmx_batch_run(dmx_batch* b)
{
    if (!b) return DMX_ERR_INVALID;
    B_HIP(hipSetDevice(b->device));
    for (;;) {
        int32_t running = 0;
        B_TRY(run_ticks(b, 2, &running));
        if (!running) break;
    }
    return DMX_OK;
}

int dmx_batch_status(const dmx_batch* b, dmx_context* c, int32_t* finished, dmx_newton_info* info)
{
    if (!b || !c) return DMX_ERR_INVALID;
    const int i = find_member(b, c);
    if (i < 0) return DMX_ERR_INVALID;
    const auto& m = b->mem[i];
    const ProbState& s = m.slot >= 0 && m.slot < b->nslots ? b->h_st[m.slot] : m.init;
    if (finished) *finished = s.status;
    if (info) {
        *info = dmx_newton_info{};
        info->converged = s.status == 1;
        info->newton_steps = s.newtonSteps;
        info->wasted_steps = s.wastedSteps;
        info->linear_iterations = s.linIts;
        info->refinements = s.refinements;
        info->time_steps = s.timeSteps;
        info->dt_retries = s.dtRetries;
        info->last_shift = s.shift;
        info->fp64_fallback = s.prec64;
    }
    return DMX_OK;
}

int dmx_batch_time_kernel(dmx_batch* b, int32_t kernel, int32_t reps, double* avg_ms, double* bytes)
{
    if (!b || reps < 1 || !avg_ms || !bytes) return DMX_ERR_INVALID;
    const bool mgKernel = kernel == DMX_KERNEL_MG_DOWN0 || kernel == DMX_KERNEL_MG_UP0;
    if (kernel != DMX_KERNEL_ASSEMBLE && kernel != DMX_KERNEL_REFINE_RESIDUAL && kernel != DMX_KERNEL_BJ_SPMV_DOT_FP32 &&
        kernel != DMX_KERNEL_BJ_SPMV_DOTS_FP32 && kernel != DMX_KERNEL_BJ_ITERATION_FP32 && !mgKernel) {
        b->err = "batch time_kernel: kernel not available in the batched executor";
        return DMX_ERR_UNSUPPORTED;
    }
    if (mgKernel && !b->prm.mg) {
        b->err = "batch time_kernel: the multigrid sweeps need DMX_PRECOND_AMG";
        return DMX_ERR_UNSUPPORTED;
    }
    const int P = b->nslots;
    if (P == 0 || b->dirty) {
        b->err = "batch time_kernel: run a batch step with the current members first";
        return DMX_ERR_INVALID;
    }
    for (int pi = 0; pi < P; ++pi)
        if (b->h_st[pi].status == 0) {
            b->err = "batch time_kernel: members still running (their Krylov vectors would be overwritten)";
            return DMX_ERR_INVALID;
        }
    B_HIP(hipSetDevice(b->device));
    std::vector<ProbState> saved(b->h_st, b->h_st + P);
    const bool krylov = kernel != DMX_KERNEL_ASSEMBLE && kernel != DMX_KERNEL_REFINE_RESIDUAL;
    if (krylov) {
        B_TRY(krylov_timing_setup(b));
    } else {
        for (int pi = 0; pi < P; ++pi)
            b->h_st[pi].phase = kernel == DMX_KERNEL_ASSEMBLE ? dmx::PH_ASM : dmx::PH_CORRECT;
        B_HIP(hipMemcpy(b->d_st, b->h_st, sizeof(ProbState) * P, hipMemcpyHostToDevice));
        B_TRY(each_kind(b, [&](auto& K) { return plan(b, K); }));
    }
    B_HIP(hipStreamSynchronize(b->stream));
    auto once = [&]() -> int {
        switch (kernel) {
        case DMX_KERNEL_ASSEMBLE: return launch_assembly(b);
        case DMX_KERNEL_REFINE_RESIDUAL:
            return each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_t_residual, K, b->d_probs, b->d_st, work_of(K), b->rw); });
        case DMX_KERNEL_BJ_SPMV_DOT_FP32:
        case DMX_KERNEL_BJ_SPMV_DOTS_FP32: {
            const int which = kernel == DMX_KERNEL_BJ_SPMV_DOT_FP32 ? 0 : 1;
            return each_kind(b, [&](auto& K) { return KLAUNCH_SPMV(dmx::k_t_spmv, K, b->d_probs, b->d_st, work_of(K), b->rw, which); });
        }
        case DMX_KERNEL_MG_DOWN0:
        case DMX_KERNEL_MG_UP0:
            return each_kind(b, [&](auto& K) {
                return K.mgTiled > 0 ? launch_mg_sweep(b, K, 0, 0, kernel == DMX_KERNEL_MG_UP0) : hipSuccess;
            });
        default: return launch_bicgstab_iteration(b);
        }
    };
    int rc = once();  // warm-up
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.f;
    if (rc == DMX_OK && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        (void)hipEventRecord(e0, b->stream);
        for (int i = 0; i < reps && rc == DMX_OK; ++i) rc = once();
        (void)hipEventRecord(e1, b->stream);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
    } else if (rc == DMX_OK) {
        b->err = "batch time_kernel: event creation failed";
        rc = DMX_ERR_DEVICE;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    // the assembly wrote an unscaled FP32 operator: scale it as a tick does
    if (rc == DMX_OK && kernel == DMX_KERNEL_ASSEMBLE)
        rc = each_kind(b, [&](auto& K) { return KLAUNCH(dmx::k_t_begin, K, b->d_probs, b->d_st, work_of(K), b->rw); });
    B_HIP(hipStreamSynchronize(b->stream));
    B_HIP(hipMemcpy(b->h_ks, b->d_ks, sizeof(KrylovScalars) * P, hipMemcpyDeviceToHost));
    bool early = false;
    for (int pi = 0; pi < P; ++pi) early |= krylov && b->h_ks[pi].done != 0;
    // restore the states
    std::copy(saved.begin(), saved.end(), b->h_st);
    B_HIP(hipMemcpy(b->d_st, b->h_st, sizeof(ProbState) * P, hipMemcpyHostToDevice));
    if (rc != DMX_OK) return rc;
    if (early) {
        b->err = "batch time_kernel: the Krylov kernels stopped early on the timing data";
        return DMX_NUMERICAL_PROBLEM;
    }
    *avg_ms = double(ms) / reps;
    // algorithmic bytes: each input read once, each output written once (as dmx_time_kernel)
    double tot = 0.0;
    for (int pi = 0; pi < P; ++pi) {
        const dmx_context* c = b->mem[pi].ctx;
        const double N = double(c->N), m = c->m, mm = m * m, nnzb = double(c->nnzb), n = double(c->n);
        const double inst = c->havePrev ? 1.0 : 0.0;
        // scaled stencil operator: the present off-diagonal blocks, the input, the output, and
        // for B the vector rt of the fused dot (D's dots use its input and output)
        const double offd = 4 * mm * (nnzb - N);
        switch (kernel) {
        case DMX_KERNEL_ASSEMBLE: tot += 8 * N * (m + m * inst + 1) + 8 * n + (8 + 4) * mm * nnzb + 4 * mm * N; break;
        case DMX_KERNEL_REFINE_RESIDUAL: tot += nnzb * (8 * mm + 4) + 4 * (N + 1) + 3 * 8 * n; break;
        case DMX_KERNEL_BJ_SPMV_DOT_FP32: tot += offd + 3 * 4 * n; break;
        case DMX_KERNEL_BJ_SPMV_DOTS_FP32: tot += offd + 2 * 4 * n; break;
        case DMX_KERNEL_MG_DOWN0:
        case DMX_KERNEL_MG_UP0:
            if (b->h_probs[pi].mgTail > 0) tot += dmx::mg_sweep_bytes(c->g, 0, c->m, 2.0, kernel == DMX_KERNEL_MG_UP0);
            break;
        // A: r,p,v -> p; B; C: x,p,r,v -> x,r; D; E: x,r,t,rt -> x,r
        default:
            tot += 4 * 4 * n + (offd + 3 * 4 * n) + 6 * 4 * n + (offd + 2 * 4 * n) + 6 * 4 * n;
            if (b->prm.mg) tot += 2 * dmx::mg_vcycle_bytes(c->g, c->m, 2.0);   // y = M^-1 p and z = M^-1 s
        }
    }
    *bytes = tot;
    return DMX_OK;
}

int dmx_batch_set_tuning(dmx_batch* b, int32_t coarse, int32_t vec_grid, int32_t spmv_grid)
{
    if (!b || (coarse != 1 && coarse != 2 && coarse != 4 && coarse != 8) || vec_grid < 8 || spmv_grid < 8 ||
        vec_grid > (1 << 20) || spmv_grid > (1 << 20)) {
        if (b) b->err = "set_tuning: coarse must be 1, 2, 4 or 8 and grids in [8, 2^20]";
        return DMX_ERR_INVALID;
    }
    B_HIP(hipSetDevice(b->device));
    B_HIP(hipStreamSynchronize(b->stream));
    b->tune.coarse = coarse;
    b->tune.vecGrid = vec_grid;
    b->tune.spmvGrid = spmv_grid;
    return b->dirty ? DMX_OK : capture_tick(b);
}

int dmx_batch_get_tuning(const dmx_batch* b, int32_t* coarse, int32_t* vec_grid, int32_t* spmv_grid)
{
    if (!b) return DMX_ERR_INVALID;
    if (coarse) *coarse = b->tune.coarse;
    if (vec_grid) *vec_grid = int32_t(b->tune.vecGrid);
    if (spmv_grid) *spmv_grid = int32_t(b->tune.spmvGrid);
    return DMX_OK;
}

int dmx_batch_autotune(dmx_batch* b, int32_t reps, double* best_us)
{
    if (!b || reps < 1) return DMX_ERR_INVALID;
    const int P = b->nslots;
    if (P == 0 || b->dirty) {
        b->err = "autotune: run a batch step with the current members first";
        return DMX_ERR_INVALID;
    }
    for (int pi = 0; pi < P; ++pi)
        if (b->h_st[pi].status == 0) {
            b->err = "autotune: members still running (their Krylov vectors would be overwritten)";
            return DMX_ERR_INVALID;
        }
    B_HIP(hipSetDevice(b->device));
    const std::vector<ProbState> saved(b->h_st, b->h_st + P);
    const dmx_batch::Tune keep = b->tune;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    B_HIP(hipEventCreate(&e0));
    B_HIP(hipEventCreate(&e1));
    int rc = DMX_OK;
    // ms per iteration of one shape on the same fresh solve (two warm-up iterations), or a
    // negative value when a member's solve stopped early (less work: not comparable)
    auto measure = [&](const dmx_batch::Tune& t, int n) -> double {
        b->tune = t;
        if (rc == DMX_OK) rc = krylov_timing_setup(b);
        for (int i = 0; i < 2 && rc == DMX_OK; ++i) rc = launch_bicgstab_iteration(b);
        if (rc != DMX_OK) return -1.0;
        (void)hipEventRecord(e0, b->stream);
        for (int i = 0; i < n && rc == DMX_OK; ++i) rc = launch_bicgstab_iteration(b);
        (void)hipEventRecord(e1, b->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipMemcpy(b->h_ks, b->d_ks, sizeof(KrylovScalars) * P, hipMemcpyDeviceToHost);
        for (int pi = 0; pi < P; ++pi)
            if (b->h_ks[pi].done) return -1.0;
        return double(ms) / n;
    };
    // pass 1: every shape once
    std::vector<std::pair<double, dmx_batch::Tune>> cand;
    for (int coarse : {1, 2, 4, 8})
        for (int64_t vg : {1024, 2048, 4096, 8192})
            for (int64_t sg : {2048, 4096, 8192, 16384}) {
                dmx_batch::Tune t;
                t.coarse = coarse;
                t.vecGrid = vg;
                t.spmvGrid = sg;
                const double ms = measure(t, reps);
                if (ms > 0.0) cand.push_back({ms, t});
            }
    std::sort(cand.begin(), cand.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    // pass 2: the four fastest and the current shape, three interleaved rounds of 3*reps
    // iterations each, ranked by their median; the current shape stays unless beaten by > 2 %
    std::vector<dmx_batch::Tune> finals{keep};
    for (size_t i = 0; i < cand.size() && finals.size() < 5; ++i) finals.push_back(cand[i].second);
    std::vector<std::vector<double>> times(finals.size());
    for (int round = 0; round < 3 && rc == DMX_OK; ++round)
        for (size_t i = 0; i < finals.size(); ++i) times[i].push_back(measure(finals[i], 3 * reps));
    double bestMs = 1e300, keepMs = 1e300;
    dmx_batch::Tune best = keep;
    for (size_t i = 0; i < finals.size(); ++i) {
        std::sort(times[i].begin(), times[i].end());
        const double med = times[i].size() == 3 && times[i][0] > 0.0 ? times[i][1] : 1e300;
        if (i == 0) keepMs = med;
        if (med < bestMs) {
            bestMs = med;
            best = finals[i];
        }
    }
    if (keepMs < 1e299 && bestMs > 0.98 * keepMs) {
        best = keep;
        bestMs = keepMs;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::copy(saved.begin(), saved.end(), b->h_st);
    B_HIP(hipMemcpy(b->d_st, b->h_st, sizeof(ProbState) * P, hipMemcpyHostToDevice));
    b->tune = bestMs < 1e299 ? best : keep;
    if (rc != DMX_OK) return rc;
    if (best_us) *best_us = bestMs < 1e299 ? 1e3 * bestMs : -1.0;
    return capture_tick(b);
}

} // extern "C"
