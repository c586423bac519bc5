// Aggregation multigrid as the preconditioner of the single-problem BiCGSTAB (dmx_linear_solve
// with DMX_PRECOND_AMG; the role of AMGBiCGSTABIstlSolver, linear/istlsolvers.hh:729-736).
//
// The hierarchy is mg.hh's (2x2x2 aggregates, Galerkin coarse operators accumulated in FP64 and
// stored in FP32, damped block-Jacobi smoothing) on the unscaled Jacobian: level 0 is the
// assembled J itself as an FP32 stencil operator with explicit diagonal blocks, so its residual
// is restricted by plain summation and the V-cycle approximates J^-1 directly (the batched
// executor, which iterates on D^-1 J, restricts by P^T D instead). The V-cycle runs the
// batched executor's box-tiled sweeps and one-workgroup tail (mg_sweeps.hh) on this hierarchy:
// a launch per level and direction down to the tail level, then the tail.
#include "kernels.hh"
#include "mg_sweeps.hh"

namespace dmx {

namespace {

constexpr int kThreads = 256;

inline dim3 grid_of(int64_t n) { return dim3(unsigned((n + kThreads - 1) / kThreads)); }

__device__ __forceinline__ bool mg_done(const KrylovScalars* ks) { return ks && ks->done; }

// level 0 from the FP64 BCRS Jacobian: off-diagonal blocks into the stencil layout (zeros where
// a direction has no neighbour), diagonal blocks and their inverses (inverted in FP64)
template <int M>
__global__ void __launch_bounds__(kThreads) k_mg1_level0(GridDesc g, const int32_t* __restrict__ rowptr,
                                                         const double* __restrict__ J, MgLevel L, int32_t* fail)
{
    constexpr int MM = M * M;
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= g.N) return;
    int idx[3];
    mg_coords(g, i, idx);
    const int64_t row0 = rowptr[i];
    auto pos = [&](int ax, int dir) { return g.dim == 3 ? stencil_pos<3>(g, idx, ax, dir) : stencil_pos<2>(g, idx, ax, dir); };
    const double* Bd = J + (row0 + pos(0, 0)) * MM;
    double Dd[MM];
    for (int q = 0; q < MM; ++q) {
        Dd[q] = Bd[q];
        L.D[i * MM + q] = float(Bd[q]);
    }
    if (!invert_block<M>(Dd)) *fail = 1;
    for (int q = 0; q < MM; ++q) L.Dinv[i * MM + q] = float(Dd[q]);
    for (int d = 0; d < 2 * g.dim; ++d) {
        const int ax = d >> 1;
        const bool up = d & 1;
        const bool has = up ? idx[ax] < g.n[ax] - 1 : idx[ax] > 0;
        const double* B = has ? J + (row0 + pos(ax, up ? 1 : -1)) * MM : nullptr;
        for (int q = 0; q < MM; ++q) L.S[(int64_t(d) * MM + q) * g.N + i] = has ? float(B[q]) : 0.f;
    }
}

template <int M>
__global__ void __launch_bounds__(kThreads) k_mg1_galerkin_j(GridDesc fg, const int32_t* __restrict__ rowptr,
                                                             const double* __restrict__ J, float* D0, MgLevel c)
{
    const int64_t C = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (C >= c.g.N) return;
    int cidx[3];
    mg_coords(c.g, C, cidx);
    mg_galerkin_jacobian<M>(fg, rowptr, J, D0, c, cidx, C);
}

template <int M>
__global__ void __launch_bounds__(kThreads) k_mg1_galerkin(MgLevel f, MgLevel c)
{
    const int64_t C = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (C >= c.g.N) return;
    int cidx[3];
    mg_coords(c.g, C, cidx);
    mg_galerkin<M>(f, c, cidx, C);
}

// down-sweep of level l over its boxes (lev: the hierarchy in device memory)
template <int M, typename R>
__global__ void __launch_bounds__(kMgBoxCells) k_mg1_down(const MgLevel* __restrict__ lev, int l, const R* rhs,
                                                          const KrylovScalars* ks)
{
    if (mg_done(ks)) return;
    mg_down_box<M>(lev[l], lev[l + 1], rhs, int32_t(blockIdx.x));
}

template <int M, typename R, typename O>
__global__ void __launch_bounds__(kMgTailThreads) k_mg1_tail(const MgLevel* __restrict__ lev, int nl, int lt,
                                                             const R* rhs, O* out, const KrylovScalars* ks)
{
    __shared__ float lev_lds[kMgTailLdsCells * mg_tail_row<M>()];
    __shared__ float r_lds[kMgTailCells * M];
    if (mg_done(ks)) return;
    mg_tail_cycle<M>(lev, nl, lt, rhs, out, lev_lds, r_lds);
}

template <int M, typename R, typename O>
__global__ void __launch_bounds__(kMgBoxCells) k_mg1_up(const MgLevel* __restrict__ lev, int l, const R* rhs,
                                                        O* out, const KrylovScalars* ks)
{
    if (mg_done(ks)) return;
    mg_up_box<M>(lev[l], lev[l + 1].g, lev[l + 1].f, rhs, out, int32_t(blockIdx.x));
}

} // namespace

hipError_t launch_mg1_setup(hipStream_t s, int m, const GridDesc& g, const int32_t* rowptr, const double* J,
                            const MgLevel* lev, int nl, int32_t* fail)
{
#define DMX_MG1(MM)                                                                                  \
    hipLaunchKernelGGL((k_mg1_level0<MM>), grid_of(g.N), dim3(kThreads), 0, s, g, rowptr, J, lev[0], fail); \
    for (int l = 1; l < nl; ++l) {                                                                   \
        if (l == 1)                                                                                  \
            hipLaunchKernelGGL((k_mg1_galerkin_j<MM>), grid_of(lev[1].g.N), dim3(kThreads), 0, s, g, rowptr, J, lev[0].D, lev[1]); \
        else                                                                                         \
            hipLaunchKernelGGL((k_mg1_galerkin<MM>), grid_of(lev[l].g.N), dim3(kThreads), 0, s, lev[l - 1], lev[l]); \
    }
    if (m == 1) { DMX_MG1(1) } else { DMX_MG1(2) }
#undef DMX_MG1
    return hipGetLastError();
}

template <typename T>
hipError_t launch_mg1_apply(hipStream_t s, int m, const MgLevel* lev, const MgLevel* levDev, int nl, int lt, const T* in,
                            T* out, const KrylovScalars* ks)
{
    // level 0 reads the operand and writes the result in T (FP64 rounded to FP32 on load and
    // widened on store: the bits of converting around an FP32 cycle); the levels below are FP32
#define DMX_MG1(MM)                                                                                   \
    for (int l = 0; l < lt; ++l) {                                                                    \
        const dim3 g(unsigned(lev[l].box.ntiles)), b(kMgBoxCells);                                    \
        if (l == 0) hipLaunchKernelGGL((k_mg1_down<MM, T>), g, b, 0, s, levDev, 0, in, ks);           \
        else hipLaunchKernelGGL((k_mg1_down<MM, float>), g, b, 0, s, levDev, l, lev[l].f, ks);       \
    }                                                                                                 \
    if (lt == 0)                                                                                      \
        hipLaunchKernelGGL((k_mg1_tail<MM, T, T>), dim3(1), dim3(kMgTailThreads), 0, s, levDev, nl, 0, in, out, ks); \
    else                                                                                              \
        hipLaunchKernelGGL((k_mg1_tail<MM, float, float>), dim3(1), dim3(kMgTailThreads), 0, s, levDev, nl, lt, \
                           lev[lt].f, lev[lt].f, ks);                                                 \
    for (int l = lt - 1; l >= 0; --l) {                                                               \
        const dim3 g(unsigned(lev[l].box.ntiles)), b(kMgBoxCells);                                    \
        if (l == 0) hipLaunchKernelGGL((k_mg1_up<MM, T, T>), g, b, 0, s, levDev, 0, in, out, ks);     \
        else hipLaunchKernelGGL((k_mg1_up<MM, float, float>), g, b, 0, s, levDev, l, lev[l].f, lev[l].f, ks); \
    }
    if (m == 1) { DMX_MG1(1) } else { DMX_MG1(2) }
#undef DMX_MG1
    return hipGetLastError();
}

template hipError_t launch_mg1_apply<float>(hipStream_t, int, const MgLevel*, const MgLevel*, int, int, const float*, float*,
                                            const KrylovScalars*);
template hipError_t launch_mg1_apply<double>(hipStream_t, int, const MgLevel*, const MgLevel*, int, int, const double*,
                                             double*, const KrylovScalars*);

} // namespace dmx
