// gfx950 assembly kernels of the CCTpfa Newton path.
//
//  k_cell_state   : phase pressures and rho*lambda of every cell (undeflected volume variables)
//  k_ghost_state  : Dirichlet ghost states of the boundary faces
//  k_assemble     : one thread per cell I: residual r[I] and, with JAC, either the
//                   forward-difference column I of the Jacobian (A[I][I] and A[J][I] for the
//                   face neighbours J) or, with AN, the hand-coded derivatives of row I
//                   (A[I][I] and A[I][J])
// The per-cell arithmetic lives in cellops.hh and is shared with no host code path of the
// library; it restates assembly/cclocalassembler.hh:164-348 in the reference operation order.
//
// Blocks are remapped so that each of the 8 XCDs owns a contiguous range of cells: the
// hardware dispatches workgroups round-robin over XCDs, and contiguous ranges keep the
// neighbour reads of a workgroup inside one XCD's L2.
#include "kernels.hh"

namespace dmx {

namespace {

constexpr int kXcds = 8;

__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nb)
{
    const int64_t per = (nb + kXcds - 1) / kXcds;
    return (b % kXcds) * per + b / kXcds;
}

inline int64_t xcd_grid(int64_t nb)
{
    const int64_t per = (nb + kXcds - 1) / kXcds;
    return per * kXcds;
}

template <int M>
__global__ void __launch_bounds__(256) k_cell_state(Phys ph, const double* __restrict__ u, double* __restrict__ vvp,
                                                    double* __restrict__ vvu, double* __restrict__ vvd,
                                                    const int32_t* __restrict__ mat, int64_t N, int64_t nb)
{
    const int64_t blk = xcd_remap(blockIdx.x, nb);
    if (blk >= nb) return;
    const int64_t I = blk * blockDim.x + threadIdx.x;
    if (I >= N) return;
    cell_state<M>(ph, u, vvp, vvu, vvd, mat, I);
}

template <int M>
__global__ void __launch_bounds__(256) k_ghost_state(Phys ph, const int32_t* __restrict__ bct, const double* __restrict__ bcv,
                                                     const int32_t* __restrict__ bmat,
                                                     double* __restrict__ gp, double* __restrict__ gu, int64_t nbf)
{
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= nbf) return;
    ghost_state<M>(ph, bct, bcv, bmat, gp, gu, b);
}

template <int DIM, int M, bool JAC, bool AN>
__global__ void __launch_bounds__(128) k_assemble(GridDesc g, Phys ph, AsmPtrs P, int64_t nb)
{
    const int64_t blk = xcd_remap(blockIdx.x, nb);
    if (blk >= nb) return;
    const int64_t I = blk * blockDim.x + threadIdx.x;
    if (I >= g.N) return;
    if constexpr (JAC && !AN) {
        // partial reassembly (cclocalassembler.hh:61-76): a green cell keeps its column of the
        // previous Jacobian and evaluates its residual only
        if (P.green && P.green[I]) {
            assemble_cell<DIM, M, false, false>(g, ph, P, I);
            return;
        }
    }
    assemble_cell<DIM, M, JAC, AN>(g, ph, P, I);
}

} // namespace

hipError_t launch_cell_state(hipStream_t s, const Phys& ph, const double* u, double* vvp, double* vvu, double* vvd,
                             const int32_t* mat, int64_t N)
{
    if (N <= 0) return hipSuccess;
    const int64_t nb = (N + 255) / 256;
    if (ph.m == 1)
        hipLaunchKernelGGL(k_cell_state<1>, dim3(xcd_grid(nb)), dim3(256), 0, s, ph, u, vvp, vvu, nullptr, mat, N, nb);
    else
        hipLaunchKernelGGL(k_cell_state<2>, dim3(xcd_grid(nb)), dim3(256), 0, s, ph, u, vvp, vvu, vvd, mat, N, nb);
    return hipGetLastError();
}

hipError_t launch_ghost_state(hipStream_t s, const Phys& ph, const int32_t* bct, const double* bcv, const int32_t* bmat,
                              double* gp, double* gu, int64_t nbf)
{
    if (nbf <= 0) return hipSuccess;
    const int64_t nb = (nbf + 255) / 256;
    if (ph.m == 1)
        hipLaunchKernelGGL(k_ghost_state<1>, dim3(nb), dim3(256), 0, s, ph, bct, bcv, bmat, gp, gu, nbf);
    else
        hipLaunchKernelGGL(k_ghost_state<2>, dim3(nb), dim3(256), 0, s, ph, bct, bcv, bmat, gp, gu, nbf);
    return hipGetLastError();
}

hipError_t launch_assemble(hipStream_t s, const GridDesc& g, const Phys& ph, const AsmPtrs& P, bool jacobian)
{
    if (g.N <= 0) return hipSuccess;
    const int threads = 128;
    const int64_t nb = (g.N + threads - 1) / threads;
    const dim3 grid(xcd_grid(nb)), block(threads);
    if (jacobian && P.analytic && ph.m == 2 && !P.vv_d) return hipErrorInvalidValue;
#define DMX_ASM_CASE(D, MM)                                                                              \
    if (g.dim == D && ph.m == MM) {                                                                      \
        if (!jacobian) hipLaunchKernelGGL((k_assemble<D, MM, false, false>), grid, block, 0, s, g, ph, P, nb); \
        else if (P.analytic) hipLaunchKernelGGL((k_assemble<D, MM, true, true>), grid, block, 0, s, g, ph, P, nb); \
        else hipLaunchKernelGGL((k_assemble<D, MM, true, false>), grid, block, 0, s, g, ph, P, nb);     \
        return hipGetLastError();                                                                        \
    }
    DMX_ASM_CASE(2, 1)
    DMX_ASM_CASE(2, 2)
    DMX_ASM_CASE(3, 1)
    DMX_ASM_CASE(3, 2)
#undef DMX_ASM_CASE
    return hipErrorInvalidValue;
}

} // namespace dmx
