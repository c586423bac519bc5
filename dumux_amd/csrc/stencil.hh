// Structured-stencil FP32 operator of the batched Krylov solver.
//
// The FP64 Jacobian stays in the reference's BCRS layout. The FP32 operator used inside the
// iterative refinement is an internal copy laid out for the GPU: the grid is structured, so the
// column of every off-diagonal block follows from the cell index and the direction, and no
// column indices are stored. Off-diagonal blocks are kept structure-of-arrays,
//     S[(d * M*M + q) * N + i],   d = 2*axis + (1 on the + side), q = a*M + b (row a, column b),
// block (d, i) = dr_i / du_j with j the neighbour of cell i in direction d; directions without a
// neighbour hold zeros. After scale_row the operator is left-scaled by the inverted diagonal
// blocks (block Jacobi as a left preconditioner): S_i <- D_i^-1 S_i, so the scaled diagonal is
// the identity and y_i = x_i + sum_d S(d, i) x_j.
#pragma once
#include "cellops.hh"

namespace dmx {

DMX_HD void cell_coords(const GridDesc& g, int64_t i, int (&idx)[3])
{
    const int64_t n01 = int64_t(g.n[0]) * g.n[1];
    idx[2] = int(i / n01);
    const int64_t r = i - int64_t(idx[2]) * n01;
    idx[1] = int(r / g.n[0]);
    idx[0] = int(r - int64_t(idx[1]) * g.n[0]);
}

// neighbour of the cell with coordinates idx in direction d (false: boundary)
DMX_HD bool stencil_neighbor(const GridDesc& g, const int (&idx)[3], int64_t i, int d, int64_t& j)
{
    const int ax = d >> 1;
    if (d & 1) {
        if (idx[ax] >= g.n[ax] - 1) return false;
        j = i + g.stride[ax];
    } else {
        if (idx[ax] <= 0) return false;
        j = i - g.stride[ax];
    }
    return true;
}

// S(., i) <- D_i^-1 S(., i) for present directions, zeros for absent ones
template <int M>
DMX_HD void scale_row(const GridDesc& g, float* S, const float* Dinv, int64_t i)
{
    int idx[3];
    cell_coords(g, i, idx);
    const int64_t N = g.N;
    float D[M * M];
    for (int q = 0; q < M * M; ++q) D[q] = Dinv[i * M * M + q];
    for (int d = 0; d < 2 * g.dim; ++d) {
        int64_t j;
        const bool has = stencil_neighbor(g, idx, i, d, j);
        float B[M * M];
        for (int q = 0; q < M * M; ++q) B[q] = has ? S[(int64_t(d) * M * M + q) * N + i] : 0.f;
        for (int a = 0; a < M; ++a)
            for (int b = 0; b < M; ++b) {
                float s = 0.f;
                for (int c = 0; c < M; ++c) s += D[a * M + c] * B[c * M + b];
                S[(int64_t(d) * M * M + a * M + b) * N + i] = s;
            }
    }
}

// y = x_i + sum_d S(d, i) x_neighbour  (scaled operator, identity diagonal)
template <int M>
DMX_HD void stencil_apply(const GridDesc& g, const float* S, const float* x, int64_t i, float (&y)[M])
{
    int idx[3];
    cell_coords(g, i, idx);
    const int64_t N = g.N;
    for (int a = 0; a < M; ++a) y[a] = x[i * M + a];
    for (int d = 0; d < 2 * g.dim; ++d) {
        int64_t j;
        if (!stencil_neighbor(g, idx, i, d, j)) continue;
        float xj[M];
        for (int b = 0; b < M; ++b) xj[b] = x[j * M + b];
        for (int a = 0; a < M; ++a) {
            float s = 0.f;
            for (int b = 0; b < M; ++b) s += S[(int64_t(d) * M * M + a * M + b) * N + i] * xj[b];
            y[a] += s;
        }
    }
}

} // namespace dmx
