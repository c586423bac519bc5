// gfx950 block-sparse linear algebra for the Newton linear systems.
//
// BCRS with m x m row-major blocks (m = 1 or 2), int32 indices. Everything is HBM-bound: no
// MFMA. The BiCGSTAB driver keeps its scalars (rho, alpha, omega, norms) in device memory
// (KrylovScalars); the kernels read them directly and turn into no-ops once the solve has
// converged or broken down, so the host only polls a flag every few iterations.
//
// Reductions are deterministic: a fixed grid of at most REDUCE_MAX_BLOCKS workgroups, each
// summing a fixed contiguous chunk (wave64 butterfly + 4 waves in fixed order) into a partial,
// then a one-workgroup finalize kernel summing the partials in index order and applying the
// scalar step (no device-scope fences inside a kernel).
//
// ILU(0) follows dune-istl's block ILU0 (third-party, called from linear/istlsolvers.hh:635-642).
// The factorization runs level by level (rows of a level are independent). The two triangular
// sweeps run either level by level or, on the structured stencil, as one launch each over boxes
// of cells (below). Within a row the operations run in the sequential order, so the factors and
// the preconditioned vectors equal the sequential ones bit for bit in every form.
#include "kernels.hh"
#include "krylov_ops.hh"

namespace dmx {

namespace {

constexpr int kThreads = 256;
constexpr double kBreakdownEps = 1e-80;

inline int reduce_blocks(int64_t n)
{
    int64_t b = (n + kThreads * 4 - 1) / (kThreads * 4);
    if (b < 1) b = 1;
    if (b > REDUCE_MAX_BLOCKS) b = REDUCE_MAX_BLOCKS;
    return int(b);
}

// wave64 xor butterfly (all lanes end with the same bits), then the 4 waves in fixed order
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* lds)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    const T r = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_max(double v, double* lds)
{
    for (int o = 32; o > 0; o >>= 1) v = __builtin_fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = __builtin_fmax(__builtin_fmax(lds[0], lds[1]), __builtin_fmax(lds[2], lds[3]));
    __syncthreads();
    return r;
}

// this workgroup's NV partial sums -> partial[k*cap + blockIdx.x]
template <int NV>
__device__ __forceinline__ void store_partials(const double* vals, ReduceWork w)
{
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) w.partial[size_t(k) * w.cap + blockIdx.x] = vals[k];
}

// ---------------------------------------------------------------------------------------------
// residual r = b - A x (row per lane)
// ---------------------------------------------------------------------------------------------
template <int M>
__global__ void __launch_bounds__(kThreads) k_residual(int64_t nrows, const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ colidx, const double* __restrict__ A,
                                                       const double* __restrict__ x, const double* __restrict__ b,
                                                       double* __restrict__ r)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= nrows) return;
    double acc[M];
    for (int a = 0; a < M; ++a) acc[a] = 0.0;
    const int32_t e = rowptr[i + 1];
    for (int32_t k = rowptr[i]; k < e; ++k) {
        const int64_t j = colidx[k];
        const double* B = A + int64_t(k) * M * M;
        for (int a = 0; a < M; ++a)
            for (int c = 0; c < M; ++c) acc[a] += B[a * M + c] * x[j * M + c];
    }
    for (int a = 0; a < M; ++a) r[i * M + a] = b[i * M + a] - acc[a];
}

// ---------------------------------------------------------------------------------------------
// ILU(0)
// ---------------------------------------------------------------------------------------------
template <int M>
__device__ __forceinline__ bool blk_invert(double* a)
{
    if (M == 1) {
        if (__builtin_fabs(a[0]) <= kBreakdownEps) return false;
        a[0] = 1.0 / a[0];
        return true;
    }
    double detinv = a[0] * a[3] - a[1] * a[2];
    if (__builtin_fabs(detinv) <= kBreakdownEps) return false;
    detinv = 1.0 / detinv;
    const double temp = a[0];
    a[0] = a[3] * detinv;
    a[1] = -a[1] * detinv;
    a[2] = -a[2] * detinv;
    a[3] = temp * detinv;
    return true;
}

template <int M>
__global__ void __launch_bounds__(kThreads) k_ilu0_factor_level(const int32_t* __restrict__ rowptr,
                                                                const int32_t* __restrict__ colidx,
                                                                const int32_t* __restrict__ diag, double* LU,
                                                                const int32_t* __restrict__ rows, int32_t count,
                                                                int32_t* fail)
{
    const int32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= count) return;
    constexpr int MM = M * M;
    const int64_t i = rows[t];
    const int32_t endi = rowptr[i + 1];
    int32_t ij = rowptr[i];
    for (; ij < endi && colidx[ij] < i; ++ij) {
        const int64_t k = colidx[ij];
        const int32_t kk = diag[k];
        // L_ik = A_ik * inv(A_kk)  (rightmultiply)
        double C[MM], L[MM];
        for (int q = 0; q < MM; ++q) C[q] = LU[int64_t(ij) * MM + q];
        for (int a = 0; a < M; ++a)
            for (int b = 0; b < M; ++b) {
                double s = 0;
                for (int c = 0; c < M; ++c) s += C[a * M + c] * LU[int64_t(kk) * MM + c * M + b];
                L[a * M + b] = s;
            }
        for (int q = 0; q < MM; ++q) LU[int64_t(ij) * MM + q] = L[q];
        int32_t jk = kk + 1;
        const int32_t endk = rowptr[k + 1];
        int32_t ik = ij + 1;
        while (ik != endi && jk != endk) {
            const int32_t ci = colidx[ik], cj = colidx[jk];
            if (ci == cj) {
                // B = L_ik * A_kj (leftmultiply), A_ij -= B
                double Bk[MM];
                for (int q = 0; q < MM; ++q) Bk[q] = LU[int64_t(jk) * MM + q];
                for (int a = 0; a < M; ++a)
                    for (int b = 0; b < M; ++b) {
                        double s = 0;
                        for (int c = 0; c < M; ++c) s += L[a * M + c] * Bk[c * M + b];
                        LU[int64_t(ik) * MM + a * M + b] -= s;
                    }
                ++ik;
                ++jk;
            } else if (ci < cj) {
                ++ik;
            } else {
                ++jk;
            }
        }
    }
    if (ij == endi || colidx[ij] != i) { atomicOr(fail, 1); return; }
    double D[MM];
    for (int q = 0; q < MM; ++q) D[q] = LU[int64_t(ij) * MM + q];
    if (!blk_invert<M>(D)) atomicOr(fail, 1);
    for (int q = 0; q < MM; ++q) LU[int64_t(ij) * MM + q] = D[q];
}

template <int M, typename T>
__global__ void __launch_bounds__(kThreads) k_ilu0_lower(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                                         const T* __restrict__ LU, const int32_t* __restrict__ rows,
                                                         int32_t count, const T* __restrict__ d, T* v,
                                                         const KrylovScalars* ks)
{
    if (ks && ks->done) return;
    const int32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= count) return;
    const int64_t i = rows[t];
    T rhs[M];
    for (int a = 0; a < M; ++a) rhs[a] = d[i * M + a];
    const int32_t e = rowptr[i + 1];
    for (int32_t k = rowptr[i]; k < e; ++k) {
        const int64_t j = colidx[k];
        if (j >= i) break;
        const T* A = LU + int64_t(k) * M * M;
        for (int a = 0; a < M; ++a)
            for (int b = 0; b < M; ++b) rhs[a] -= A[a * M + b] * v[j * M + b];
    }
    for (int a = 0; a < M; ++a) v[i * M + a] = rhs[a];
}

template <int M, typename T>
__global__ void __launch_bounds__(kThreads) k_ilu0_upper(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                                         const int32_t* __restrict__ diag, const T* __restrict__ LU,
                                                         const int32_t* __restrict__ rows, int32_t count, T* v,
                                                         const KrylovScalars* ks)
{
    if (ks && ks->done) return;
    const int32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= count) return;
    const int64_t i = rows[t];
    T rhs[M];
    for (int a = 0; a < M; ++a) rhs[a] = v[i * M + a];
    const int32_t dk = diag[i];
    for (int32_t k = rowptr[i + 1] - 1; k > dk; --k) {
        const int64_t j = colidx[k];
        const T* A = LU + int64_t(k) * M * M;
        for (int a = 0; a < M; ++a)
            for (int b = 0; b < M; ++b) rhs[a] -= A[a * M + b] * v[j * M + b];
    }
    const T* D = LU + int64_t(dk) * M * M;
    for (int a = 0; a < M; ++a) {
        T s = T(0);
        for (int b = 0; b < M; ++b) s += D[a * M + b] * rhs[b];
        v[i * M + a] = s;
    }
}

// ---------------------------------------------------------------------------------------------
// ILU(0) sweeps in one launch each, by boxes of cells
//
// The structured grid is cut into boxes (TileGeom: at most kTileCells cells). A workgroup solves
// one box at a time in LDS:
//   1. it loads the box rows' right-hand sides and triangle blocks;
//   2. it waits for the halo, the solved values of the cells just outside the box's lower faces
//      (upper faces in the backward sweep), which other workgroups produce;
//   3. it runs the box's own levels lx+ly+lz = 0, 1, ... with a barrier between levels;
//   4. it publishes the box's values.
// Boxes are drawn from a counter in the order of their box level tx+ty+tz (reversed for the
// backward sweep). The counter only grows, so every box being worked on has all boxes it waits
// for drawn by running workgroups, independent of the dispatch order. The waits are on the
// values themselves: the output vector starts filled with a signalling-NaN pattern that
// arithmetic never produces, values are stored and polled with relaxed device-scope atomics,
// and a halo cell is ready once it no longer holds the pattern. Every wait is bounded: after
// kSpinLimit polls the box is published as quiet NaNs and fail |= 2.
//
// Per row the arithmetic is the level kernels' above, operand for operand (blocks in column
// order, then the inverted diagonal block in the backward sweep), so both forms give the same
// bits; only the synchronization differs.
// ---------------------------------------------------------------------------------------------
template <typename T> struct SweepMark;
template <> struct SweepMark<double> {
    using U = uint64_t;
    static constexpr U mark = 0x7FF4A5A5A5A5A5A5ull;   // signalling NaN
    static constexpr U qnan = 0x7FF8000000000000ull;
};
template <> struct SweepMark<float> {
    using U = uint32_t;
    static constexpr U mark = 0x7FA5A5A5u;
    static constexpr U qnan = 0x7FC00000u;
};

constexpr int kSweepSide = kIluSweepMaxSide;   // off-diagonal blocks per triangle of a row
constexpr uint32_t kSpinLimit = 1u << 20;

template <typename T>
__device__ __forceinline__ T sweep_mark()
{
    return __builtin_bit_cast(T, SweepMark<T>::mark);
}

template <typename T>
__device__ __forceinline__ void sweep_publish(T* p, T v, bool valid)
{
    using U = typename SweepMark<T>::U;
    U bits = __builtin_bit_cast(U, v);
    if (!valid || bits == SweepMark<T>::mark) bits = SweepMark<T>::qnan;
    __hip_atomic_store(reinterpret_cast<U*>(p), bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// polls p until it no longer holds the pattern; false after kSpinLimit polls
template <typename T>
__device__ __forceinline__ bool sweep_wait(const T* p, T& out)
{
    using U = typename SweepMark<T>::U;
    const U* q = reinterpret_cast<const U*>(p);
    U bits = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t spins = 0;
    while (bits == SweepMark<T>::mark && spins < kSpinLimit) {
        __builtin_amdgcn_s_sleep(1);
        bits = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ++spins;
    }
    out = __builtin_bit_cast(T, bits);
    return bits != SweepMark<T>::mark;
}

// BACK = false: v = L^-1 in (unit lower triangle), every box cell also sets reset[i] to the
//               pattern (readying the backward sweep's output).
// BACK = true:  v = U^-1 in (upper triangle, inverted diagonal blocks); in[i] is read once and
//               set back to the pattern for the next forward sweep.
// tickets[0] / tickets[1] count the forward / backward sweep's boxes; each sweep clears the
// other's counter for the launch that follows it.
// NT threads per workgroup, Q box cells per thread (NT * Q >= box cells)
template <int M, typename T, bool BACK, int NT, int Q>
__global__ void __launch_bounds__(NT) k_ilu0_box_sweep(GridDesc g, TileGeom tg, const int32_t* __restrict__ order,
                                                             const int32_t* __restrict__ rowptr,
                                                             const int32_t* __restrict__ colidx,
                                                             const int32_t* __restrict__ diag, const T* __restrict__ LU,
                                                             T* in, T* v, T* __restrict__ reset, uint32_t* tickets,
                                                             int32_t* fail, const KrylovScalars* ks)
{
    if (ks && ks->done) return;
    constexpr int MM = M * M;
    constexpr int SH = BACK ? 0 : 1;                  // halo at local -1 (forward) or at the extent
    __shared__ T xs[kTileLds * M];
    __shared__ int32_t s_box, s_bad;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(&tickets[BACK ? 0 : 1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int px = tg.b[0] + 1, py = tg.b[1] + 1;
    const int64_t s1 = g.stride[1], s2 = g.stride[2];
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            s_box = int32_t(atomicAdd(&tickets[BACK ? 1 : 0], 1u));
            s_bad = 0;
        }
        __syncthreads();
        const int32_t pos = s_box;
        if (pos >= tg.ntiles) break;
        const int32_t box = order[BACK ? tg.ntiles - 1 - pos : pos];
        const int bx = box % tg.nt[0], by = (box / tg.nt[0]) % tg.nt[1], bz = box / (tg.nt[0] * tg.nt[1]);
        const int o0 = bx * tg.b[0], o1 = by * tg.b[1], o2 = bz * tg.b[2];
        const int e0 = min(tg.b[0], g.n[0] - o0), e1 = min(tg.b[1], g.n[1] - o1), e2 = min(tg.b[2], g.n[2] - o2);
        const int ncell = e0 * e1 * e2;

        // 1. the rows of this thread's cells
        int lvl[Q], own[Q], nb[Q];
        int64_t gi[Q];
        int nbx[Q][kSweepSide];
        T B[Q][kSweepSide][MM], D[Q][MM], rhs[Q][M];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int c = threadIdx.x + q * NT;
            lvl[q] = -1;
            nb[q] = 0;
            own[q] = 0;
            gi[q] = 0;
            if (c >= ncell) continue;
            const int lx = c % e0, ly = (c / e0) % e1, lz = c / (e0 * e1);
            const int64_t i = int64_t(o2 + lz) * s2 + int64_t(o1 + ly) * s1 + (o0 + lx);
            gi[q] = i;
            own[q] = ((lz + SH) * py + (ly + SH)) * px + (lx + SH);
            lvl[q] = BACK ? (e0 - 1 - lx) + (e1 - 1 - ly) + (e2 - 1 - lz) : lx + ly + lz;
            for (int a = 0; a < M; ++a) {
                rhs[q][a] = in[i * M + a];
                if (BACK) in[i * M + a] = sweep_mark<T>();
                else reset[i * M + a] = sweep_mark<T>();
            }
            const int32_t dk = diag[i];
            const int32_t kb = BACK ? rowptr[i + 1] - 1 : rowptr[i];
            const int32_t cnt = BACK ? rowptr[i + 1] - 1 - dk : dk - rowptr[i];
#pragma unroll
            for (int l = 0; l < kSweepSide; ++l) {
                if (l >= cnt) break;
                const int32_t k = BACK ? kb - l : kb + l;
                const int64_t j = colidx[k];
                const int64_t jz = j / s2, jr = j - jz * s2, jy = jr / s1, jx = jr - jy * s1;
                nbx[q][l] = ((int(jz) - o2 + SH) * py + (int(jy) - o1 + SH)) * px + (int(jx) - o0 + SH);
                for (int t = 0; t < MM; ++t) B[q][l][t] = LU[int64_t(k) * MM + t];
                nb[q] = l + 1;
            }
            if (BACK)
                for (int t = 0; t < MM; ++t) D[q][t] = LU[int64_t(dk) * MM + t];
        }

        // 2. the halo: faces of the neighbouring boxes that come earlier in this sweep
        const bool h0 = BACK ? o0 + e0 < g.n[0] : o0 > 0;
        const bool h1 = BACK ? o1 + e1 < g.n[1] : o1 > 0;
        const bool h2 = BACK ? o2 + e2 < g.n[2] : o2 > 0;
        const int f0 = h0 ? e1 * e2 : 0, f1 = h1 ? e0 * e2 : 0, f2 = h2 ? e0 * e1 : 0;
        for (int h = threadIdx.x; h < f0 + f1 + f2; h += NT) {
            int lx, ly, lz;
            if (h < f0) {
                ly = h % e1; lz = h / e1; lx = BACK ? e0 : -1;
            } else if (h < f0 + f1) {
                const int r = h - f0;
                lx = r % e0; lz = r / e0; ly = BACK ? e1 : -1;
            } else {
                const int r = h - f0 - f1;
                lx = r % e0; ly = r / e0; lz = BACK ? e2 : -1;
            }
            const int64_t j = int64_t(o2 + lz) * s2 + int64_t(o1 + ly) * s1 + (o0 + lx);
            const int slot = ((lz + SH) * py + (ly + SH)) * px + (lx + SH);
            for (int b = 0; b < M; ++b) {
                T x;
                if (!sweep_wait(&v[j * M + b], x)) {
                    s_bad = 1;
                    atomicOr(fail, 2);
                }
                xs[slot * M + b] = x;
            }
        }
        __syncthreads();

        // 3. the box's levels
        const int nlev = e0 + e1 + e2 - 2;
        for (int L = 0; L < nlev; ++L) {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                if (lvl[q] != L) continue;
                T x[M];
                for (int a = 0; a < M; ++a) x[a] = rhs[q][a];
#pragma unroll
                for (int l = 0; l < kSweepSide; ++l) {
                    if (l >= nb[q]) break;
                    for (int a = 0; a < M; ++a)
                        for (int b = 0; b < M; ++b) x[a] -= B[q][l][a * M + b] * xs[nbx[q][l] * M + b];
                }
                if (BACK) {
                    for (int a = 0; a < M; ++a) {
                        T sum = T(0);
                        for (int b = 0; b < M; ++b) sum += D[q][a * M + b] * x[b];
                        xs[own[q] * M + a] = sum;
                    }
                } else {
                    for (int a = 0; a < M; ++a) xs[own[q] * M + a] = x[a];
                }
            }
            __syncthreads();
        }

        // 4. publish
        const bool ok = s_bad == 0;
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (lvl[q] >= 0)
                for (int a = 0; a < M; ++a) sweep_publish(&v[gi[q] * M + a], xs[own[q] * M + a], ok);
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_fill_mark(int64_t n, T* x)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    const SweepMark<T> sm{};
    if (i < n) x[i] = __builtin_bit_cast(T, sm.mark);
}

// ---------------------------------------------------------------------------------------------
// block Jacobi
// ---------------------------------------------------------------------------------------------
template <int M, typename T>
__global__ void __launch_bounds__(kThreads) k_jacobi_setup(int64_t nrows, const int32_t* __restrict__ diag,
                                                           const double* __restrict__ A, T* __restrict__ Dinv, int32_t* fail)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= nrows) return;
    double D[M * M];
    for (int q = 0; q < M * M; ++q) D[q] = A[int64_t(diag[i]) * M * M + q];
    if (!blk_invert<M>(D)) atomicOr(fail, 1);
    for (int q = 0; q < M * M; ++q) Dinv[i * M * M + q] = T(D[q]);
}

template <int M, typename T>
__global__ void __launch_bounds__(kThreads) k_jacobi_apply(int64_t nrows, const T* __restrict__ Dinv, const T* __restrict__ p,
                                                           T* __restrict__ y, const KrylovScalars* ks)
{
    if (ks && ks->done) return;
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= nrows) return;
    T pv[M];
    for (int b = 0; b < M; ++b) pv[b] = p[i * M + b];
    for (int a = 0; a < M; ++a) {
        T s = T(0);
        for (int b = 0; b < M; ++b) s += Dinv[i * M * M + a * M + b] * pv[b];
        y[i * M + a] = s;
    }
}

// ---------------------------------------------------------------------------------------------
// BiCGSTAB pieces (dune-istl BiCGSTABSolver::apply step order)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void chunk_range(int64_t n, int64_t& beg, int64_t& end)
{
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    beg = int64_t(blockIdx.x) * chunk;
    end = beg + chunk < n ? beg + chunk : n;
}

__device__ __forceinline__ bool finite(double x) { return __builtin_isfinite(x); }

template <typename T>
__global__ void __launch_bounds__(kThreads) k_bicg_init(int64_t n, const double* __restrict__ b, T* r, T* rt, T* p, T* v, T* x,
                                                        ReduceWork w, double bscale)
{
    __shared__ T lds[4];
    int64_t beg, end;
    chunk_range(n, beg, end);
    T acc = T(0);
    for (int64_t i = beg + threadIdx.x; i < end; i += kThreads) {
        const T ri = T(b[i] * bscale);
        r[i] = ri;
        rt[i] = ri;
        p[i] = T(0);
        v[i] = T(0);
        x[i] = T(0);
        acc += ri * ri;
    }
    const double part = double(block_sum<T>(acc, lds));
    store_partials<1>(&part, w);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_bicg_update_p(int64_t n, const T* __restrict__ r, const T* __restrict__ v,
                                                            T* __restrict__ p, int first, const KrylovScalars* ks)
{
    if (ks->done) return;
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n) return;
    if (first) { p[i] = r[i]; return; }
    const double beta = (ks->norm == 0.0) ? 0.0 : (ks->rho_new / ks->rho) * (ks->alpha / ks->omega);
    const T om = T(ks->omega), bt = T(beta);
    T pi = p[i];
    pi += -om * v[i];
    pi *= bt;
    pi += r[i];
    p[i] = pi;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_bicg_half1(int64_t n, const T* __restrict__ y, const T* __restrict__ v,
                                                         T* __restrict__ x, T* __restrict__ r, const KrylovScalars* ks, ReduceWork w)
{
    if (ks->done) return;
    __shared__ T lds[4];
    const T al = T(ks->alpha);
    int64_t beg, end;
    chunk_range(n, beg, end);
    T acc = T(0);
    for (int64_t i = beg + threadIdx.x; i < end; i += kThreads) {
        x[i] += al * y[i];
        const T ri = r[i] + (-al) * v[i];
        r[i] = ri;
        acc += ri * ri;
    }
    const double part = double(block_sum<T>(acc, lds));
    store_partials<1>(&part, w);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_bicg_half2(int64_t n, const T* __restrict__ y, const T* __restrict__ t,
                                                         const T* __restrict__ rt, T* __restrict__ x, T* __restrict__ r,
                                                         const KrylovScalars* ks, ReduceWork w)
{
    if (ks->done) return;
    __shared__ T lds[4];
    const T om = T(ks->omega);
    int64_t beg, end;
    chunk_range(n, beg, end);
    T a1 = T(0), a2 = T(0);
    for (int64_t i = beg + threadIdx.x; i < end; i += kThreads) {
        x[i] += om * y[i];
        const T ri = r[i] + (-om) * t[i];
        r[i] = ri;
        a1 += ri * ri;
        a2 += rt[i] * ri;
    }
    double part[2];
    part[0] = double(block_sum<T>(a1, lds));
    part[1] = double(block_sum<T>(a2, lds));
    store_partials<2>(part, w);
}

// one workgroup: sum (or max) the partials in index order, then the scalar logic of `op`
__global__ void __launch_bounds__(kThreads) k_finalize(int op, int64_t nparts, KrylovScalars* ks, ReduceWork w,
                                                       double* out, double reduction, int32_t maxit)
{
    __shared__ double lds[4];
    if (ks && op != FIN_INIT && op != FIN_NORM && op != FIN_MAX && op != FIN_SUM && ks->done) return;
    const int nv = (op == FIN_OMEGA || op == FIN_HALF2) ? 2 : 1;
    double tot[2] = {0.0, 0.0};
    for (int k = 0; k < nv; ++k) {
        if (op == FIN_MAX) {
            double acc = 0.0;
            for (int64_t b = threadIdx.x; b < nparts; b += kThreads) acc = __builtin_fmax(acc, w.partial[size_t(k) * w.cap + b]);
            tot[k] = block_max(acc, lds);
        } else {
            double acc = 0.0;
            for (int64_t b = threadIdx.x; b < nparts; b += kThreads) acc += w.partial[size_t(k) * w.cap + b];
            tot[k] = block_sum<double>(acc, lds);
        }
    }
    if (threadIdx.x != 0) return;
    fin_apply(op, tot, ks, out, reduction, maxit);
}

// ---------------------------------------------------------------------------------------------
// Newton update + shift, norms, conversions
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) k_newton_update(int64_t n, const double* __restrict__ uLast,
                                                            const double* __restrict__ du, double* __restrict__ u, ReduceWork w)
{
    __shared__ double lds[4];
    int64_t beg, end;
    chunk_range(n, beg, end);
    double mx = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kThreads) {
        const double u1 = uLast[i];
        double u2 = u1;
        u2 += -1.0 * du[i];
        u[i] = u2;
        const double d = __builtin_fabs(u1 - u2) / __builtin_fmax(1.0, __builtin_fabs(u1 + u2) * 0.5);
        mx = __builtin_fmax(mx, d);
    }
    const double part = block_max(mx, lds);
    store_partials<1>(&part, w);
}

// Partial reassembly after a Newton update (newtonsolver.hh:565-603 with
// PartialReassemblerEngine<CCTpfa>::computeColors, partialreassembler.hh:308-346): cell i's
// distance from the last linearization grows by the maximum relative shift of its unknowns
// between uLast and uLast - (uLast - u) (the update as newtonUpdate forms it, the max in
// std::max's order); cells above the threshold are red (reassembled, distance reset), the
// others green. Partials: green cells per block.
__global__ void __launch_bounds__(kThreads) k_partial_colors(int64_t ncells, int m, const double* __restrict__ uLast,
                                                             const double* __restrict__ u, double threshold,
                                                             double* __restrict__ dist, uint8_t* __restrict__ green, ReduceWork w)
{
    __shared__ double lds[4];
    int64_t beg, end;
    chunk_range(ncells, beg, end);
    double ng = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kThreads) {
        double s = 0.0;
        for (int a = 0; a < m; ++a) {
            const double cur = uLast[i * m + a];
            double delta = cur;
            delta -= u[i * m + a];
            double next = cur;
            next -= delta;
            const double d = __builtin_fabs(cur - next) / __builtin_fmax(1.0, __builtin_fabs(cur + next) * 0.5);
            s = s < d ? d : s;
        }
        double di = dist[i];
        di += s;
        const bool red = di > threshold;
        green[i] = red ? 0 : 1;
        dist[i] = red ? 0.0 : di;
        ng += red ? 0.0 : 1.0;
    }
    const double part = block_sum<double>(ng, lds);
    store_partials<1>(&part, w);
}

__global__ void __launch_bounds__(kThreads) k_norm(int64_t n, const double* __restrict__ x, ReduceWork w)
{
    __shared__ double lds[4];
    int64_t beg, end;
    chunk_range(n, beg, end);
    double acc = 0.0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kThreads) acc += x[i] * x[i];
    const double part = block_sum<double>(acc, lds);
    store_partials<1>(&part, w);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_add_to(int64_t n, const T* __restrict__ y, double scale, double* __restrict__ x)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) x[i] += scale * double(y[i]);
}

__global__ void __launch_bounds__(kThreads) k_to_f32(int64_t n, const double* __restrict__ x, float* __restrict__ y)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) y[i] = float(x[i]);
}

__global__ void __launch_bounds__(kThreads) k_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) y[i] = x[i];
}

__global__ void __launch_bounds__(kThreads) k_fill(int64_t n, double* __restrict__ x, double v)
{
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) x[i] = v;
}

inline dim3 grid1(int64_t n) { return dim3(unsigned((n + kThreads - 1) / kThreads)); }

} // namespace

hipError_t launch_residual(hipStream_t s, int m, int64_t nrows, const int32_t* rowptr, const int32_t* colidx,
                           const double* A, const double* x, const double* b, double* r)
{
    if (nrows <= 0) return hipSuccess;
    if (m == 1) hipLaunchKernelGGL((k_residual<1>), grid1(nrows), dim3(kThreads), 0, s, nrows, rowptr, colidx, A, x, b, r);
    else hipLaunchKernelGGL((k_residual<2>), grid1(nrows), dim3(kThreads), 0, s, nrows, rowptr, colidx, A, x, b, r);
    return hipGetLastError();
}

hipError_t launch_ilu0_factor_level(hipStream_t s, int m, const int32_t* rowptr, const int32_t* colidx,
                                    const int32_t* diag, double* LU, const int32_t* rows, int32_t count, int32_t* fail)
{
    if (count <= 0) return hipSuccess;
    if (m == 1) hipLaunchKernelGGL((k_ilu0_factor_level<1>), grid1(count), dim3(kThreads), 0, s, rowptr, colidx, diag, LU, rows, count, fail);
    else hipLaunchKernelGGL((k_ilu0_factor_level<2>), grid1(count), dim3(kThreads), 0, s, rowptr, colidx, diag, LU, rows, count, fail);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ilu0_lower_level(hipStream_t s, int m, const int32_t* rowptr, const int32_t* colidx, const T* LU,
                                   const int32_t* rows, int32_t count, const T* d, T* v, const KrylovScalars* ks)
{
    if (count <= 0) return hipSuccess;
    if (m == 1) hipLaunchKernelGGL((k_ilu0_lower<1, T>), grid1(count), dim3(kThreads), 0, s, rowptr, colidx, LU, rows, count, d, v, ks);
    else hipLaunchKernelGGL((k_ilu0_lower<2, T>), grid1(count), dim3(kThreads), 0, s, rowptr, colidx, LU, rows, count, d, v, ks);
    return hipGetLastError();
}
template hipError_t launch_ilu0_lower_level<double>(hipStream_t, int, const int32_t*, const int32_t*, const double*, const int32_t*, int32_t, const double*, double*, const KrylovScalars*);
template hipError_t launch_ilu0_lower_level<float>(hipStream_t, int, const int32_t*, const int32_t*, const float*, const int32_t*, int32_t, const float*, float*, const KrylovScalars*);

template <typename T>
hipError_t launch_ilu0_upper_level(hipStream_t s, int m, const int32_t* rowptr, const int32_t* colidx, const int32_t* diag,
                                   const T* LU, const int32_t* rows, int32_t count, T* v, const KrylovScalars* ks)
{
    if (count <= 0) return hipSuccess;
    if (m == 1) hipLaunchKernelGGL((k_ilu0_upper<1, T>), grid1(count), dim3(kThreads), 0, s, rowptr, colidx, diag, LU, rows, count, v, ks);
    else hipLaunchKernelGGL((k_ilu0_upper<2, T>), grid1(count), dim3(kThreads), 0, s, rowptr, colidx, diag, LU, rows, count, v, ks);
    return hipGetLastError();
}
template hipError_t launch_ilu0_upper_level<double>(hipStream_t, int, const int32_t*, const int32_t*, const int32_t*, const double*, const int32_t*, int32_t, double*, const KrylovScalars*);
template hipError_t launch_ilu0_upper_level<float>(hipStream_t, int, const int32_t*, const int32_t*, const int32_t*, const float*, const int32_t*, int32_t, float*, const KrylovScalars*);

template <typename T>
hipError_t launch_ilu0_sweeps(hipStream_t s, int m, const GridDesc& g, const TileGeom& tg, const int32_t* order,
                              const int32_t* rowptr, const int32_t* colidx, const int32_t* diag, const T* LU, int grid,
                              const T* d, T* tmp, T* y, uint32_t* tickets, int32_t* fail, const KrylovScalars* ks)
{
    if (tg.ntiles <= 0) return hipSuccess;
    if ((tg.b[0] + 1) * (tg.b[1] + 1) * (tg.b[2] + 1) > kTileLds || tg.b[0] * tg.b[1] * tg.b[2] > kTileCells)
        return hipErrorInvalidValue;
    const unsigned gr = unsigned(grid < tg.ntiles ? grid : tg.ntiles);
    T* dd = const_cast<T*>(d);   // the forward sweep only reads d
    const int cells = tg.b[0] * tg.b[1] * tg.b[2];
#define DMX_BOX_SWEEPS(MM_, NT_, Q_)                                                                          \
    do {                                                                                                     \
        hipLaunchKernelGGL((k_ilu0_box_sweep<MM_, T, false, NT_, Q_>), dim3(gr), dim3(NT_), 0, s, g, tg, order, \
                           rowptr, colidx, diag, LU, dd, tmp, y, tickets, fail, ks);                         \
        hipLaunchKernelGGL((k_ilu0_box_sweep<MM_, T, true, NT_, Q_>), dim3(gr), dim3(NT_), 0, s, g, tg, order,  \
                           rowptr, colidx, diag, LU, tmp, y, nullptr, tickets, fail, ks);                    \
    } while (0)
#define DMX_BOX_SWEEPS_M(NT_, Q_)                    \
    do {                                             \
        if (m == 1) DMX_BOX_SWEEPS(1, NT_, Q_);      \
        else DMX_BOX_SWEEPS(2, NT_, Q_);             \
    } while (0)
    if (cells <= 64) DMX_BOX_SWEEPS_M(64, 1);
    else if (cells <= 128) DMX_BOX_SWEEPS_M(128, 1);
    else if (cells <= 256) DMX_BOX_SWEEPS_M(256, 1);
    else DMX_BOX_SWEEPS_M(256, 2);
#undef DMX_BOX_SWEEPS_M
#undef DMX_BOX_SWEEPS
    return hipGetLastError();
}
template hipError_t launch_ilu0_sweeps<double>(hipStream_t, int, const GridDesc&, const TileGeom&, const int32_t*,
                                               const int32_t*, const int32_t*, const int32_t*, const double*, int,
                                               const double*, double*, double*, uint32_t*, int32_t*, const KrylovScalars*);
template hipError_t launch_ilu0_sweeps<float>(hipStream_t, int, const GridDesc&, const TileGeom&, const int32_t*,
                                              const int32_t*, const int32_t*, const int32_t*, const float*, int,
                                              const float*, float*, float*, uint32_t*, int32_t*, const KrylovScalars*);

template <typename T>
hipError_t launch_fill_mark(hipStream_t s, T* x, int64_t n)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_fill_mark<T>), grid1(n), dim3(kThreads), 0, s, n, x);
    return hipGetLastError();
}
template hipError_t launch_fill_mark<double>(hipStream_t, double*, int64_t);
template hipError_t launch_fill_mark<float>(hipStream_t, float*, int64_t);

template <typename T>
hipError_t launch_jacobi_setup(hipStream_t s, int m, int64_t nrows, const int32_t* diag, const double* A, T* Dinv, int32_t* fail)
{
    if (nrows <= 0) return hipSuccess;
    if (m == 1) hipLaunchKernelGGL((k_jacobi_setup<1, T>), grid1(nrows), dim3(kThreads), 0, s, nrows, diag, A, Dinv, fail);
    else hipLaunchKernelGGL((k_jacobi_setup<2, T>), grid1(nrows), dim3(kThreads), 0, s, nrows, diag, A, Dinv, fail);
    return hipGetLastError();
}
template hipError_t launch_jacobi_setup<double>(hipStream_t, int, int64_t, const int32_t*, const double*, double*, int32_t*);
template hipError_t launch_jacobi_setup<float>(hipStream_t, int, int64_t, const int32_t*, const double*, float*, int32_t*);

template <typename T>
hipError_t launch_jacobi_apply(hipStream_t s, int m, int64_t nrows, const T* Dinv, const T* p, T* y, const KrylovScalars* ks)
{
    if (nrows <= 0) return hipSuccess;
    if (m == 1) hipLaunchKernelGGL((k_jacobi_apply<1, T>), grid1(nrows), dim3(kThreads), 0, s, nrows, Dinv, p, y, ks);
    else hipLaunchKernelGGL((k_jacobi_apply<2, T>), grid1(nrows), dim3(kThreads), 0, s, nrows, Dinv, p, y, ks);
    return hipGetLastError();
}
template hipError_t launch_jacobi_apply<double>(hipStream_t, int, int64_t, const double*, const double*, double*, const KrylovScalars*);
template hipError_t launch_jacobi_apply<float>(hipStream_t, int, int64_t, const float*, const float*, float*, const KrylovScalars*);

template <typename T>
hipError_t bicg_init(hipStream_t s, int64_t n, const double* b, T* r, T* rt, T* p, T* v, T* x, KrylovScalars* ks,
                     ReduceWork w, double reduction, double bscale)
{
    const int nb = reduce_blocks(n);
    hipLaunchKernelGGL((k_bicg_init<T>), dim3(nb), dim3(kThreads), 0, s, n, b, r, rt, p, v, x, w, bscale);
    return launch_finalize(s, FIN_INIT, nb, ks, w, nullptr, reduction, 0);
}
template <typename T>
hipError_t bicg_update_p(hipStream_t s, int64_t n, const T* r, const T* v, T* p, int first, KrylovScalars* ks)
{
    hipLaunchKernelGGL((k_bicg_update_p<T>), grid1(n), dim3(kThreads), 0, s, n, r, v, p, first, ks);
    return hipGetLastError();
}
template <typename T>
hipError_t bicg_half1(hipStream_t s, int64_t n, const T* y, const T* v, T* x, T* r, KrylovScalars* ks, ReduceWork w)
{
    const int nb = reduce_blocks(n);
    hipLaunchKernelGGL((k_bicg_half1<T>), dim3(nb), dim3(kThreads), 0, s, n, y, v, x, r, ks, w);
    return launch_finalize(s, FIN_HALF1, nb, ks, w, nullptr, 0.0, 0);
}
template <typename T>
hipError_t bicg_half2(hipStream_t s, int64_t n, const T* y, const T* t, const T* rt, T* x, T* r, KrylovScalars* ks,
                      ReduceWork w, int32_t maxit)
{
    const int nb = reduce_blocks(n);
    hipLaunchKernelGGL((k_bicg_half2<T>), dim3(nb), dim3(kThreads), 0, s, n, y, t, rt, x, r, ks, w);
    return launch_finalize(s, FIN_HALF2, nb, ks, w, nullptr, 0.0, maxit);
}

hipError_t launch_finalize(hipStream_t s, int op, int64_t nparts, KrylovScalars* ks, ReduceWork w, double* out,
                           double reduction, int32_t maxit)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kThreads), 0, s, op, nparts, ks, w, out, reduction, maxit);
    return hipGetLastError();
}

#define DMX_BICG_INST(T)                                                                                        \
    template hipError_t bicg_init<T>(hipStream_t, int64_t, const double*, T*, T*, T*, T*, T*, KrylovScalars*, ReduceWork, double, double); \
    template hipError_t bicg_update_p<T>(hipStream_t, int64_t, const T*, const T*, T*, int, KrylovScalars*);         \
    template hipError_t bicg_half1<T>(hipStream_t, int64_t, const T*, const T*, T*, T*, KrylovScalars*, ReduceWork); \
    template hipError_t bicg_half2<T>(hipStream_t, int64_t, const T*, const T*, const T*, T*, T*, KrylovScalars*, ReduceWork, int32_t);
DMX_BICG_INST(double)
DMX_BICG_INST(float)
#undef DMX_BICG_INST

hipError_t launch_newton_update(hipStream_t s, int64_t n, const double* uLast, const double* du, double* u,
                                double* shift_out, ReduceWork w)
{
    const int nb = reduce_blocks(n);
    hipLaunchKernelGGL(k_newton_update, dim3(nb), dim3(kThreads), 0, s, n, uLast, du, u, w);
    return launch_finalize(s, FIN_MAX, nb, nullptr, w, shift_out, 0.0, 0);
}

hipError_t launch_partial_colors(hipStream_t s, int64_t ncells, int m, const double* uLast, const double* u, double threshold,
                                 double* dist, uint8_t* green, double* ngreen_out, ReduceWork w)
{
    const int nb = reduce_blocks(ncells);
    hipLaunchKernelGGL(k_partial_colors, dim3(nb), dim3(kThreads), 0, s, ncells, m, uLast, u, threshold, dist, green, w);
    return launch_finalize(s, FIN_SUM, nb, nullptr, w, ngreen_out, 0.0, 0);
}

hipError_t launch_norm(hipStream_t s, int64_t n, const double* x, double* out, ReduceWork w)
{
    const int nb = reduce_blocks(n);
    hipLaunchKernelGGL(k_norm, dim3(nb), dim3(kThreads), 0, s, n, x, w);
    return launch_finalize(s, FIN_NORM, nb, nullptr, w, out, 0.0, 0);
}

template <typename T>
hipError_t launch_add_to(hipStream_t s, int64_t n, const T* y, double scale, double* x)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_add_to<T>), grid1(n), dim3(kThreads), 0, s, n, y, scale, x);
    return hipGetLastError();
}
template hipError_t launch_add_to<float>(hipStream_t, int64_t, const float*, double, double*);
template hipError_t launch_add_to<double>(hipStream_t, int64_t, const double*, double, double*);

hipError_t launch_to_f32(hipStream_t s, int64_t n, const double* r, float* r32)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_f32, grid1(n), dim3(kThreads), 0, s, n, r, r32);
    return hipGetLastError();
}

hipError_t launch_convert_f32(hipStream_t s, const double* x, float* y, int64_t n) { return launch_to_f32(s, n, x, y); }

hipError_t launch_copy(hipStream_t s, const double* x, double* y, int64_t n)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy, grid1(n), dim3(kThreads), 0, s, n, x, y);
    return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, double* x, double v, int64_t n)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, grid1(n), dim3(kThreads), 0, s, n, x, v);
    return hipGetLastError();
}

} // namespace dmx
