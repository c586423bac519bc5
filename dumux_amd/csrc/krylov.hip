// Fused BiCGSTAB iteration with block-Jacobi preconditioning (gfx950).
//
// One thread per cell (block row), 256 cells per workgroup. A BiCGSTAB iteration
// (dune-istl BiCGSTABSolver::apply step order, third-party) is five kernels:
//   A  p = r + beta (p - omega v);  y = Dinv p                       (row-local)
//   B  v = A y (tile SpMV);          h = <rt, v>, alpha = rho_new / h
//   C  x += alpha y; r -= alpha v;   ||r||, convergence;  y = Dinv r (row-local)
//   D  t = A y (tile SpMV);          <t, r>, <t, t>, omega
//   E  x += omega y; r -= omega t;   ||r||, <rt, r>, convergence, breakdown tests
// The scalars live in KrylovScalars (device memory); every kernel returns at entry once the
// solve has converged or broken down, so a captured HIP graph of several iterations can be
// replayed without host round trips.
//
// Tile SpMV: the blocks of the workgroup's 256 consecutive rows are contiguous in BCRS order;
// lanes take consecutive blocks (coalesced m*m values + column index), multiply with the
// gathered x block and park the m partial sums in LDS; each lane then sums its row's blocks
// from LDS in column order.
//
// Reductions: wave64 butterfly (xor shuffles: every lane holds the same bits), 4 waves via
// LDS in fixed order into one partial per workgroup; a one-workgroup finalize kernel (linalg.hip)
// sums the partials in index order and applies the scalar step. Results depend only on the
// problem size, not on timing, and no kernel needs a device-scope fence.
#include "kernels.hh"

namespace dmx {

namespace {

constexpr int kT = 256;              // threads = rows per workgroup
constexpr int kMaxRowBlocks = 7;     // 2*dim+1 point stencil

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sum over the workgroup; every thread gets the result
template <typename T>
__device__ __forceinline__ T wg_sum(T v, T* lds4)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    const T r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    __syncthreads();
    return r;
}

template <int NV>
__device__ __forceinline__ void store_partials(const double* part, ReduceWork w)
{
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) w.partial[size_t(k) * w.cap + blockIdx.x] = part[k];
}

// y_row = (A x)_row for the workgroup's rows; returns false for lanes past the last row
template <int M, typename T>
__device__ __forceinline__ bool tile_spmv(int64_t nrows, const int32_t* __restrict__ rowptr,
                                          const int32_t* __restrict__ colidx, const T* __restrict__ A,
                                          const T* __restrict__ x, T* lds, T (&acc)[M])
{
    const int64_t r0 = int64_t(blockIdx.x) * kT;
    const int64_t r1 = r0 + kT < nrows ? r0 + kT : nrows;
    const int32_t k0 = rowptr[r0], k1 = rowptr[r1];
    for (int32_t k = k0 + threadIdx.x; k < k1; k += kT) {
        const int64_t j = colidx[k];
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
    const int32_t e = rowptr[row + 1];
    for (int32_t k = rowptr[row]; k < e; ++k)
        for (int a = 0; a < M; ++a) acc[a] += lds[(k - k0) * M + a];
    return true;
}

template <int M, typename T>
__device__ __forceinline__ void jac_apply(const T* __restrict__ Dinv, int64_t i, const T (&p)[M], T (&y)[M])
{
    for (int a = 0; a < M; ++a) {
        T s = T(0);
        for (int b = 0; b < M; ++b) s += Dinv[i * M * M + a * M + b] * p[b];
        y[a] = s;
    }
}

// A: p update + preconditioner
template <int M, typename T>
__global__ void __launch_bounds__(kT) k_bj_a(int64_t nrows, const T* __restrict__ r, const T* __restrict__ v, T* __restrict__ p,
                                             T* __restrict__ y, const T* __restrict__ Dinv, const KrylovScalars* ks)
{
    if (ks->done) return;
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    if (i >= nrows) return;
    const bool first = ks->half_steps == 0;
    const double beta = (ks->norm == 0.0) ? 0.0 : (ks->rho_new / ks->rho) * (ks->alpha / ks->omega);
    const T om = T(ks->omega), bt = T(beta);
    T pv[M], yv[M];
    for (int a = 0; a < M; ++a) {
        const T ri = r[i * M + a];
        if (first) {
            pv[a] = ri;
        } else {
            T q = p[i * M + a];
            q += -om * v[i * M + a];
            q *= bt;
            q += ri;
            pv[a] = q;
        }
        p[i * M + a] = pv[a];
    }
    jac_apply<M, T>(Dinv, i, pv, yv);
    for (int a = 0; a < M; ++a) y[i * M + a] = yv[a];
}

// B: v = A y, h = <rt, v>
template <int M, typename T>
__global__ void __launch_bounds__(kT) k_bj_b(int64_t nrows, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                             const T* __restrict__ A, const T* __restrict__ y, const T* __restrict__ rt,
                                             T* __restrict__ v, const KrylovScalars* ks, ReduceWork w)
{
    if (ks->done) return;
    __shared__ T lds[kT * kMaxRowBlocks * M];
    __shared__ T lds4[4];
    T acc[M];
    T d = T(0);
    if (tile_spmv<M, T>(nrows, rowptr, colidx, A, y, lds, acc)) {
        const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
        for (int a = 0; a < M; ++a) {
            v[i * M + a] = acc[a];
            d += rt[i * M + a] * acc[a];
        }
    }
    const double part = double(wg_sum<T>(d, lds4));
    store_partials<1>(&part, w);
}

// C: x += alpha y; r -= alpha v; ||r||; y = Dinv r
template <int M, typename T>
__global__ void __launch_bounds__(kT) k_bj_c(int64_t nrows, T* __restrict__ x, T* __restrict__ r, T* __restrict__ y,
                                             const T* __restrict__ v, const T* __restrict__ Dinv, const KrylovScalars* ks,
                                             ReduceWork w)
{
    if (ks->done) return;
    __shared__ T lds4[4];
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    T d = T(0);
    if (i < nrows) {
        const T al = T(ks->alpha);
        T rv[M], yv[M];
        for (int a = 0; a < M; ++a) {
            x[i * M + a] += al * y[i * M + a];
            const T ri = r[i * M + a] + (-al) * v[i * M + a];
            r[i * M + a] = ri;
            rv[a] = ri;
            d += ri * ri;
        }
        jac_apply<M, T>(Dinv, i, rv, yv);
        for (int a = 0; a < M; ++a) y[i * M + a] = yv[a];
    }
    const double part = double(wg_sum<T>(d, lds4));
    store_partials<1>(&part, w);
}

// D: t = A y; <t, r>, <t, t>
template <int M, typename T>
__global__ void __launch_bounds__(kT) k_bj_d(int64_t nrows, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                             const T* __restrict__ A, const T* __restrict__ y, const T* __restrict__ r,
                                             T* __restrict__ t, const KrylovScalars* ks, ReduceWork w)
{
    if (ks->done) return;
    __shared__ T lds[kT * kMaxRowBlocks * M];
    __shared__ T lds4[4];
    T acc[M];
    T d1 = T(0), d2 = T(0);
    if (tile_spmv<M, T>(nrows, rowptr, colidx, A, y, lds, acc)) {
        const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
        for (int a = 0; a < M; ++a) {
            t[i * M + a] = acc[a];
            d1 += acc[a] * r[i * M + a];
            d2 += acc[a] * acc[a];
        }
    }
    double part[2];
    part[0] = double(wg_sum<T>(d1, lds4));
    part[1] = double(wg_sum<T>(d2, lds4));
    store_partials<2>(part, w);
}

// E: x += omega y; r -= omega t; ||r||, <rt, r>
template <int M, typename T>
__global__ void __launch_bounds__(kT) k_bj_e(int64_t nrows, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ y,
                                             const T* __restrict__ t, const T* __restrict__ rt, const KrylovScalars* ks,
                                             ReduceWork w)
{
    if (ks->done) return;
    __shared__ T lds4[4];
    const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
    T d1 = T(0), d2 = T(0);
    if (i < nrows) {
        const T om = T(ks->omega);
        for (int a = 0; a < M; ++a) {
            x[i * M + a] += om * y[i * M + a];
            const T ri = r[i * M + a] + (-om) * t[i * M + a];
            r[i * M + a] = ri;
            d1 += ri * ri;
            d2 += rt[i * M + a] * ri;
        }
    }
    double part[2];
    part[0] = double(wg_sum<T>(d1, lds4));
    part[1] = double(wg_sum<T>(d2, lds4));
    store_partials<2>(part, w);
}

// tile SpMV (stand-alone)
template <int M, typename T>
__global__ void __launch_bounds__(kT) k_spmv_tile(int64_t nrows, const int32_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ colidx, const T* __restrict__ A,
                                                  const T* __restrict__ x, T* __restrict__ y, const KrylovScalars* ks)
{
    if (ks && ks->done) return;
    __shared__ T lds[kT * kMaxRowBlocks * M];
    T acc[M];
    if (tile_spmv<M, T>(nrows, rowptr, colidx, A, x, lds, acc)) {
        const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
        for (int a = 0; a < M; ++a) y[i * M + a] = acc[a];
    }
}

inline dim3 rows_grid(int64_t nrows) { return dim3(unsigned((nrows + kT - 1) / kT)); }

} // namespace

int64_t tile_reduce_blocks(int64_t nrows) { return (nrows + kT - 1) / kT; }

template <typename T>
hipError_t bj_iteration(hipStream_t s, int m, int64_t nrows, const int32_t* rowptr, const int32_t* colidx,
                        const T* A, const T* Dinv, T* const* vec, KrylovScalars* ks, ReduceWork w, int32_t maxit)
{
    // vec: r, rt, p, v, t, y, x
    T *r = vec[0], *rt = vec[1], *p = vec[2], *v = vec[3], *t = vec[4], *y = vec[5], *x = vec[6];
    const dim3 g = rows_grid(nrows), b(kT);
    const int64_t np = g.x;
    hipError_t e;
#define DMX_BJ_ITER(MM)                                                                                         \
    hipLaunchKernelGGL((k_bj_a<MM, T>), g, b, 0, s, nrows, r, v, p, y, Dinv, ks);                                 \
    hipLaunchKernelGGL((k_bj_b<MM, T>), g, b, 0, s, nrows, rowptr, colidx, A, y, rt, v, ks, w);                   \
    if ((e = launch_finalize(s, FIN_H, np, ks, w, nullptr, 0.0, maxit)) != hipSuccess) return e;                \
    hipLaunchKernelGGL((k_bj_c<MM, T>), g, b, 0, s, nrows, x, r, y, v, Dinv, ks, w);                             \
    if ((e = launch_finalize(s, FIN_HALF1, np, ks, w, nullptr, 0.0, maxit)) != hipSuccess) return e;            \
    hipLaunchKernelGGL((k_bj_d<MM, T>), g, b, 0, s, nrows, rowptr, colidx, A, y, r, t, ks, w);                   \
    if ((e = launch_finalize(s, FIN_OMEGA, np, ks, w, nullptr, 0.0, maxit)) != hipSuccess) return e;            \
    hipLaunchKernelGGL((k_bj_e<MM, T>), g, b, 0, s, nrows, x, r, y, t, rt, ks, w);                               \
    if ((e = launch_finalize(s, FIN_HALF2, np, ks, w, nullptr, 0.0, maxit)) != hipSuccess) return e;
    if (m == 1) { DMX_BJ_ITER(1) } else { DMX_BJ_ITER(2) }
#undef DMX_BJ_ITER
    return hipGetLastError();
}
template hipError_t bj_iteration<float>(hipStream_t, int, int64_t, const int32_t*, const int32_t*, const float*, const float*,
                                        float* const*, KrylovScalars*, ReduceWork, int32_t);
template hipError_t bj_iteration<double>(hipStream_t, int, int64_t, const int32_t*, const int32_t*, const double*, const double*,
                                         double* const*, KrylovScalars*, ReduceWork, int32_t);

// one of the five kernels alone (0..4 = A..E), for timing
template <typename T>
hipError_t bj_kernel(hipStream_t s, int which, int m, int64_t nrows, const int32_t* rowptr, const int32_t* colidx,
                     const T* A, const T* Dinv, T* const* vec, KrylovScalars* ks, ReduceWork w, int32_t maxit)
{
    T *r = vec[0], *rt = vec[1], *p = vec[2], *v = vec[3], *t = vec[4], *y = vec[5], *x = vec[6];
    const dim3 g = rows_grid(nrows), b(kT);
#define DMX_BJ_ONE(MM)                                                                                       \
    switch (which) {                                                                                         \
    case 0: hipLaunchKernelGGL((k_bj_a<MM, T>), g, b, 0, s, nrows, r, v, p, y, Dinv, ks); break;            \
    case 1: hipLaunchKernelGGL((k_bj_b<MM, T>), g, b, 0, s, nrows, rowptr, colidx, A, y, rt, v, ks, w); break; \
    case 2: hipLaunchKernelGGL((k_bj_c<MM, T>), g, b, 0, s, nrows, x, r, y, v, Dinv, ks, w); break;         \
    case 3: hipLaunchKernelGGL((k_bj_d<MM, T>), g, b, 0, s, nrows, rowptr, colidx, A, y, r, t, ks, w); break; \
    case 4: hipLaunchKernelGGL((k_bj_e<MM, T>), g, b, 0, s, nrows, x, r, y, t, rt, ks, w); break;            \
    default: return hipErrorInvalidValue;                                                                    \
    }
    if (m == 1) { DMX_BJ_ONE(1) } else { DMX_BJ_ONE(2) }
#undef DMX_BJ_ONE
    return hipGetLastError();
}
template hipError_t bj_kernel<float>(hipStream_t, int, int, int64_t, const int32_t*, const int32_t*, const float*, const float*,
                                     float* const*, KrylovScalars*, ReduceWork, int32_t);
template hipError_t bj_kernel<double>(hipStream_t, int, int, int64_t, const int32_t*, const int32_t*, const double*, const double*,
                                      double* const*, KrylovScalars*, ReduceWork, int32_t);

template <typename T>
hipError_t launch_spmv_tile(hipStream_t s, int m, int64_t nrows, const int32_t* rowptr, const int32_t* colidx, const T* A,
                            const T* x, T* y, const KrylovScalars* ks)
{
    if (nrows <= 0) return hipSuccess;
    if (m == 1) hipLaunchKernelGGL((k_spmv_tile<1, T>), rows_grid(nrows), dim3(kT), 0, s, nrows, rowptr, colidx, A, x, y, ks);
    else hipLaunchKernelGGL((k_spmv_tile<2, T>), rows_grid(nrows), dim3(kT), 0, s, nrows, rowptr, colidx, A, x, y, ks);
    return hipGetLastError();
}
template hipError_t launch_spmv_tile<float>(hipStream_t, int, int64_t, const int32_t*, const int32_t*, const float*, const float*,
                                            float*, const KrylovScalars*);
template hipError_t launch_spmv_tile<double>(hipStream_t, int, int64_t, const int32_t*, const int32_t*, const double*,
                                             const double*, double*, const KrylovScalars*);

} // namespace dmx
