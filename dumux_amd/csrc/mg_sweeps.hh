// Box-tiled multigrid sweeps and the V-cycle tail as device functions over an array of levels:
// the batched executor runs them on each problem's levels, the single-problem preconditioner
// on its own hierarchy. The per-cell arithmetic is mg.hh's, so both give mg_vcycle_host's
// result. Levels below the tail level run box-tiled (a thread per fine cell of one box of
// kMgBoxCells cells); the tail levels (the first at most kMgTailCells cells in global memory,
// the coarser ones at most kMgTailLdsCells cells together, copied to LDS) run in one workgroup.
#pragma once
#include "mg.hh"

namespace dmx {

constexpr int kMgTailThreads = 512;   // a thread per cell of every tail level (a 4096-cell first level with
                                       // 1024 threads, 4 cells each, measured 13 % slower per iteration)
static_assert(kMgTailThreads >= kMgTailCells, "multigrid tail launch shape");

// down-sweep of box `box` of level L into the right-hand side of Cl, a thread per fine cell in
// aggregate-major order (kMgBoxCells threads): the 2^dim children of an aggregate are
// consecutive lanes of one wave, so their residuals are summed with lane shuffles (dz, dy, dx:
// mg_down's order; absent children add -0, the identity of IEEE addition). rhs: FP32, or an
// FP64 level-0 operand rounded to FP32 on load.
template <int M, typename R>
__device__ __forceinline__ void mg_down_box(const MgLevel& L, const MgLevel& Cl, const R* rhs, int32_t box)
{
    const int lane = threadIdx.x & 63;
    int idx[3], cidx[3];
    float r[M];
    for (int a = 0; a < M; ++a) r[a] = -0.f;
    if (mg_box_child(L.g, L.box, box, threadIdx.x, idx, cidx))
        mg_fine_residual<M>(L, rhs, idx, mg_index(L.g, idx[0], idx[1], idx[2]), r);
    const int nch = L.g.dim == 3 ? 8 : 4;
    const int base = lane & ~(nch - 1);
    float acc[M];
    for (int a = 0; a < M; ++a) acc[a] = 0.f;
    for (int c = 0; c < nch; ++c)
        for (int a = 0; a < M; ++a) acc[a] += __shfl(r[a], base + c, 64);
    if (lane == base && cidx[0] < Cl.g.n[0] && cidx[1] < Cl.g.n[1] && cidx[2] < Cl.g.n[2] &&
        int(threadIdx.x) < L.box.b[0] * L.box.b[1] * L.box.b[2]) {
        const int64_t C = mg_index(Cl.g, cidx[0], cidx[1], cidx[2]);
        for (int a = 0; a < M; ++a) Cl.f[C * M + a] = acc[a];
    }
}

// up-sweep of box `box` of level L (a thread per fine cell): coarse correction e on grid cg,
// post-smoothing against rhs, result into out (FP32, or widened exactly to FP64)
template <int M, typename R, typename O>
__device__ __forceinline__ void mg_up_box(const MgLevel& L, const GridDesc& cg, const float* e, const R* rhs,
                                          O* out, int32_t box)
{
    int idx[3];
    if (!mg_box_cell(L.g, L.box, box, threadIdx.x, idx)) return;
    const int64_t i = mg_index(L.g, idx[0], idx[1], idx[2]);
    mg_up<M>(L, rhs, cg, e, idx, i, out);
}

// LDS copy of tail level l > lt (carved in level order from base)
template <int M>
__device__ __forceinline__ MgLevel mg_tail_lds(const MgLevel* lev, int lt, int l, float* base)
{
    float* q = base;
    for (int k = lt + 1; k < l; ++k) q += mg_tail_floats(lev[k].g, M);
    MgLevel L = lev[l];
    const int64_t N = L.g.N;
    L.S = q;
    q += int64_t(2) * L.g.dim * M * M * N;
    L.D = q;
    q += M * M * N;
    L.Dinv = q;
    q += M * M * N;
    L.x = q;
    q += M * N;
    L.f = q;
    L.Dr = nullptr;
    return L;
}

// LDS floats per coarse cell of the tail (6 directions)
template <int M>
constexpr int mg_tail_row() { return 8 * M * M + 2 * M; }

// The tail of the V-cycle in one workgroup of kMgTailThreads threads: level lt (input rhs,
// output out) and the coarser levels, whose operators are copied to lev_lds
// (kMgTailLdsCells * mg_tail_row<M>() floats) and whose vectors live there; r_lds holds
// kMgTailCells * M floats. Down, the coarsest sweeps, and up again; barriers order the phases,
// and global memory written by the workgroup is visible to it after a barrier. rhs and out may
// be FP64 when lt is level 0 (rounded on load and widened on store, as in the sweeps).
template <int M, typename R, typename O>
__device__ __forceinline__ void mg_tail_cycle(const MgLevel* lev, int nl, int lt, const R* rhs, O* out,
                                              float* lev_lds, float* r_lds)
{
    const int tid = threadIdx.x;
    // the coarser levels' operators into LDS
    for (int l = lt + 1; l < nl; ++l) {
        const MgLevel& G = lev[l];
        const MgLevel L = mg_tail_lds<M>(lev, lt, l, lev_lds);
        const int64_t ns = int64_t(2) * G.g.dim * M * M * G.g.N, nm = M * M * G.g.N;
        for (int64_t k = tid; k < ns; k += kMgTailThreads) L.S[k] = G.S[k];
        for (int64_t k = tid; k < nm; k += kMgTailThreads) {
            L.D[k] = G.D[k];
            L.Dinv[k] = G.Dinv[k];
        }
    }
    const MgLevel& T = lev[lt];
    // residual of level lt (its pre-smoothed x stays in global memory for the up-sweep)
    if (lt + 1 < nl)
        for (int i = tid; i < T.g.N; i += kMgTailThreads) {
            int idx[3];
            mg_coords(T.g, i, idx);
            float r[M];
            mg_fine_residual<M>(T, rhs, idx, i, r);
            for (int a = 0; a < M; ++a) r_lds[i * M + a] = r[a];
        }
    __syncthreads();
    for (int l = lt; l + 1 < nl; ++l) {
        const MgLevel F = l == lt ? T : mg_tail_lds<M>(lev, lt, l, lev_lds);
        if (l > lt) {
            if (tid < F.g.N) {
                int idx[3];
                mg_coords(F.g, tid, idx);
                float r[M];
                mg_fine_residual<M>(F, F.f, idx, tid, r);
                for (int a = 0; a < M; ++a) r_lds[tid * M + a] = r[a];
            }
            __syncthreads();
        }
        const MgLevel Cl = mg_tail_lds<M>(lev, lt, l + 1, lev_lds);
        if (tid < Cl.g.N) {
            int cidx[3];
            mg_coords(Cl.g, tid, cidx);
            float acc[M];
            for (int a = 0; a < M; ++a) acc[a] = 0.f;
            for (int dz = 0; dz < 2; ++dz)
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        const int x = 2 * cidx[0] + dx, y = 2 * cidx[1] + dy, z = 2 * cidx[2] + dz;
                        if (x >= F.g.n[0] || y >= F.g.n[1] || z >= F.g.n[2]) continue;
                        const int64_t i = mg_index(F.g, x, y, z);
                        for (int a = 0; a < M; ++a) acc[a] += r_lds[i * M + a];
                    }
            for (int a = 0; a < M; ++a) Cl.f[int64_t(tid) * M + a] = acc[a];
        }
        __syncthreads();
    }
    {
        // coarsest: the iterate exchanged through r_lds; the result replaces its f
        const MgLevel C = nl - 1 > lt ? mg_tail_lds<M>(lev, lt, nl - 1, lev_lds) : lev[nl - 1];
        const bool own = tid < C.g.N;
        float xi[M];
        if (own) {
            mg_presmooth_value<M>(C, C.f, tid, xi);
            for (int a = 0; a < M; ++a) r_lds[tid * M + a] = xi[a];
        }
        __syncthreads();
        for (int sweep = 1; sweep < kMgCoarseSweeps; ++sweep) {
            if (own) mg_coarse_update<M>(C, r_lds, tid, xi);
            __syncthreads();
            if (own)
                for (int a = 0; a < M; ++a) r_lds[tid * M + a] = xi[a];
            __syncthreads();
        }
        if (own)
            for (int a = 0; a < M; ++a) C.f[int64_t(tid) * M + a] = xi[a];
        __syncthreads();
    }
    for (int l = nl - 2; l > lt; --l) {
        const MgLevel F = mg_tail_lds<M>(lev, lt, l, lev_lds);
        const MgLevel Cl = mg_tail_lds<M>(lev, lt, l + 1, lev_lds);
        if (tid < F.g.N) {
            int idx[3];
            mg_coords(F.g, tid, idx);
            mg_up<M>(F, F.f, Cl.g, Cl.f, idx, tid, F.f);
        }
        __syncthreads();
    }
    if (lt + 1 < nl) {
        const MgLevel Cl = mg_tail_lds<M>(lev, lt, lt + 1, lev_lds);
        for (int i = tid; i < T.g.N; i += kMgTailThreads) {
            int idx[3];
            mg_coords(T.g, i, idx);
            mg_up<M>(T, rhs, Cl.g, Cl.f, idx, i, out);
        }
    }
}

} // namespace dmx
