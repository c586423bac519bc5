// Aggregation multigrid preconditioner of the batched FP32 BiCGSTAB (host + device).
//
// The reference offers algebraic multigrid as a BiCGSTAB preconditioner
// (AMGBiCGSTABIstlSolver, linear/istlsolvers.hh:729-736; dune-istl's aggregation AMG,
// third-party). On the structured grid the aggregates are fixed boxes of 2x2x2 cells (2x2 in
// 2D, fewer at the grid's end), so every level is again a 2*dim+1 point stencil of m x m blocks
// and keeps the batch operator's layout:
//   S_l[(d*M*M + q)*N_l + i]  off-diagonal blocks (d = 2*axis + side), D_l, D_l^-1 per cell.
// Level 0 is the batch operator itself, left-scaled by the block diagonal (unit diagonal).
// Coarse operators are Galerkin products with piecewise-constant prolongation (unsmoothed
// aggregation, P^T A P):
//   D_{l+1}[C]   = sum over children c of (D_l[c] + the blocks of c's faces inside C)
//   S_{l+1,d}[C] = sum over the children on C's d-face of S_{l,d}[c]
// A V-cycle: one damped block-Jacobi pre-smoothing step from zero, the residual restricted by
// summation, the coarse correction prolongated by injection, one post-smoothing step; the
// coarsest level (at most kMgCoarseCells cells) gets kMgCoarseSweeps damped block-Jacobi sweeps.
// Everything is FP32 inside the FP32 inner solve and in a fixed order (deterministic).
#pragma once
#include "stencil.hh"

namespace dmx {

constexpr int kMgMaxLevels = 12;
constexpr int kMgCoarseCells = 8;
constexpr int kMgCoarseSweeps = 8;
constexpr float kMgOmega = 0.7f;

struct MgLevel {
    GridDesc g;       // dim, n, stride, N of the level (geometry fields unused)
    float* S;         // 2*dim*m*m*N off-diagonal blocks
    float* D;         // m*m*N diagonal blocks (null on level 0: unit diagonal)
    float* Dinv;      // m*m*N inverse diagonal blocks (null on level 0)
    float* x;         // m*N pre-smoothed iterate
    float* f;         // m*N right-hand side (levels >= 1; the result of the up-sweep replaces it)
};

// grid of the next coarser level: every axis halved (rounded up)
DMX_HD GridDesc mg_coarse_grid(const GridDesc& f)
{
    GridDesc c{};
    c.dim = f.dim;
    for (int a = 0; a < 3; ++a) c.n[a] = (f.n[a] + 1) / 2;
    c.N = int64_t(c.n[0]) * c.n[1] * c.n[2];
    c.stride[0] = 1;
    c.stride[1] = c.n[0];
    c.stride[2] = int64_t(c.n[0]) * c.n[1];
    return c;
}

// levels of a grid: coarsen while the level has more than kMgCoarseCells cells (1, i.e. no
// multigrid, if kMgMaxLevels levels do not get there)
DMX_HD int mg_num_levels(const GridDesc& g)
{
    int nl = 1;
    GridDesc c = g;
    while (c.N > kMgCoarseCells && nl < kMgMaxLevels) {
        c = mg_coarse_grid(c);
        ++nl;
    }
    return c.N <= kMgCoarseCells ? nl : 1;
}

// FP32 values of the coarse levels of a grid (S, D, D^-1, x, f per level)
inline int64_t mg_floats(const GridDesc& g, int m)
{
    const int nl = mg_num_levels(g);
    GridDesc c = g;
    int64_t t = 0;
    for (int l = 1; l < nl; ++l) {
        c = mg_coarse_grid(c);
        t += (int64_t(2) * c.dim * m * m + 2 * m * m + 2 * m) * c.N;
    }
    return t;
}

// cell index of coordinates on a level
DMX_HD int64_t mg_index(const GridDesc& g, int x, int y, int z)
{
    return int64_t(z) * g.stride[2] + int64_t(y) * g.stride[1] + x;
}

// children of coarse cell C (fine coordinates 2X+dx, ...), in the order dz, dy, dx; count returned
DMX_HD int mg_children(const GridDesc& fine, const GridDesc& coarse, int64_t C, int64_t (&child)[8])
{
    int idx[3];
    cell_coords(coarse, C, idx);
    int k = 0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int x = 2 * idx[0] + dx, y = 2 * idx[1] + dy, z = 2 * idx[2] + dz;
                if (x < fine.n[0] && y < fine.n[1] && z < fine.n[2]) child[k++] = mg_index(fine, x, y, z);
            }
    return k;
}

// aggregate (coarse cell) of fine cell i
DMX_HD int64_t mg_parent(const GridDesc& fine, const GridDesc& coarse, int64_t i)
{
    int idx[3];
    cell_coords(fine, i, idx);
    return mg_index(coarse, idx[0] / 2, idx[1] / 2, idx[2] / 2);
}

template <int M>
DMX_HD void mg_invert(const float* D, float* Dinv)
{
    if (M == 1) {
        Dinv[0] = float(1.0 / double(D[0]));
        return;
    }
    const double a = D[0], b = D[1 % (M * M)], c = D[2 % (M * M)], d = D[M * M - 1];
    const double det = a * d - b * c;
    const double inv = 1.0 / det;
    Dinv[0] = float(d * inv);
    Dinv[1 % (M * M)] = float(-b * inv);
    Dinv[2 % (M * M)] = float(-c * inv);
    Dinv[M * M - 1] = float(a * inv);
}

// Galerkin coarse blocks of coarse cell C from the finer level f (unit diagonal if f.D is null)
template <int M>
DMX_HD void mg_galerkin(const MgLevel& f, const MgLevel& c, int64_t C)
{
    constexpr int MM = M * M;
    const int nd = 2 * f.g.dim;
    int64_t child[8];
    const int nc = mg_children(f.g, c.g, C, child);
    float Dc[MM];
    for (int q = 0; q < MM; ++q) Dc[q] = 0.f;
    float Sc[6][MM];
    for (int d = 0; d < nd; ++d)
        for (int q = 0; q < MM; ++q) Sc[d][q] = 0.f;
    for (int k = 0; k < nc; ++k) {
        const int64_t i = child[k];
        for (int q = 0; q < MM; ++q) Dc[q] += f.D ? f.D[i * MM + q] : ((q % (M + 1)) == 0 ? 1.f : 0.f);
        int idx[3];
        cell_coords(f.g, i, idx);
        for (int d = 0; d < nd; ++d) {
            int64_t j;
            if (!stencil_neighbor(f.g, idx, i, d, j)) continue;
            const bool inside = mg_parent(f.g, c.g, j) == C;
            for (int q = 0; q < MM; ++q) {
                const float s = f.S[(int64_t(d) * MM + q) * f.g.N + i];
                if (inside) Dc[q] += s;
                else Sc[d][q] += s;
            }
        }
    }
    for (int q = 0; q < MM; ++q) c.D[C * MM + q] = Dc[q];
    mg_invert<M>(Dc, c.Dinv + C * MM);
    for (int d = 0; d < nd; ++d)
        for (int q = 0; q < MM; ++q) c.S[(int64_t(d) * MM + q) * c.g.N + C] = Sc[d][q];
}

// y = A_l x at cell i (explicit or unit diagonal)
template <int M>
DMX_HD void mg_apply_row(const MgLevel& L, const float* x, int64_t i, float (&y)[M])
{
    constexpr int MM = M * M;
    for (int a = 0; a < M; ++a) {
        float s = 0.f;
        if (L.D)
            for (int b = 0; b < M; ++b) s += L.D[i * MM + a * M + b] * x[i * M + b];
        else
            s = x[i * M + a];
        y[a] = s;
    }
    int idx[3];
    cell_coords(L.g, i, idx);
    for (int d = 0; d < 2 * L.g.dim; ++d) {
        int64_t j;
        if (!stencil_neighbor(L.g, idx, i, d, j)) continue;
        for (int a = 0; a < M; ++a) {
            float s = 0.f;
            for (int b = 0; b < M; ++b) s += L.S[(int64_t(d) * MM + a * M + b) * L.g.N + i] * x[j * M + b];
            y[a] += s;
        }
    }
}

// z = omega D^-1 v (unit diagonal: omega v)
template <int M>
DMX_HD void mg_jacobi(const MgLevel& L, int64_t i, const float (&v)[M], float (&z)[M])
{
    for (int a = 0; a < M; ++a) {
        float s = 0.f;
        if (L.Dinv)
            for (int b = 0; b < M; ++b) s += L.Dinv[i * M * M + a * M + b] * v[b];
        else
            s = v[a];
        z[a] = kMgOmega * s;
    }
}

// pre-smoothing value of fine cell i: x = omega D^-1 f
template <int M>
DMX_HD void mg_presmooth_value(const MgLevel& L, const float* f, int64_t i, float (&x)[M])
{
    float v[M];
    for (int a = 0; a < M; ++a) v[a] = f[i * M + a];
    mg_jacobi<M>(L, i, v, x);
}

// Down-sweep at coarse cell C of level l+1: pre-smooth the children (x_l = omega D^-1 f_l, kept
// in L.x), then f_{l+1}[C] = sum over the children of (f_l - A_l x_l). The neighbours' x_l
// are recomputed from f_l (the first smoothing step is pointwise).
template <int M>
DMX_HD void mg_down(const MgLevel& L, const float* fl, const MgLevel& Cl, int64_t C)
{
    constexpr int MM = M * M;
    int64_t child[8];
    const int nc = mg_children(L.g, Cl.g, C, child);
    float acc[M];
    for (int a = 0; a < M; ++a) acc[a] = 0.f;
    for (int k = 0; k < nc; ++k) {
        const int64_t i = child[k];
        float xi[M];
        mg_presmooth_value<M>(L, fl, i, xi);
        for (int a = 0; a < M; ++a) L.x[i * M + a] = xi[a];
        float r[M];
        for (int a = 0; a < M; ++a) {
            float s = 0.f;
            if (L.D)
                for (int b = 0; b < M; ++b) s += L.D[i * MM + a * M + b] * xi[b];
            else
                s = xi[a];
            r[a] = fl[i * M + a] - s;
        }
        int idx[3];
        cell_coords(L.g, i, idx);
        for (int d = 0; d < 2 * L.g.dim; ++d) {
            int64_t j;
            if (!stencil_neighbor(L.g, idx, i, d, j)) continue;
            float xj[M];
            mg_presmooth_value<M>(L, fl, j, xj);
            for (int a = 0; a < M; ++a) {
                float s = 0.f;
                for (int b = 0; b < M; ++b) s += L.S[(int64_t(d) * MM + a * M + b) * L.g.N + i] * xj[b];
                r[a] -= s;
            }
        }
        for (int a = 0; a < M; ++a) acc[a] += r[a];
    }
    for (int a = 0; a < M; ++a) Cl.f[C * M + a] = acc[a];
}

// Up-sweep at fine cell i of level l: x = x_l + e[parent] (coarse correction, e = the coarse
// level's result), then out[i] = x_i + omega D^-1 (f_l - A_l x)_i with the corrected neighbours.
template <int M>
DMX_HD void mg_up(const MgLevel& L, const float* fl, const MgLevel& Cl, const float* e, int64_t i, float* out)
{
    constexpr int MM = M * M;
    auto corrected = [&](int64_t c, float (&v)[M]) {
        const int64_t P = mg_parent(L.g, Cl.g, c);
        for (int a = 0; a < M; ++a) v[a] = L.x[c * M + a] + e[P * M + a];
    };
    float xi[M];
    corrected(i, xi);
    float r[M];
    for (int a = 0; a < M; ++a) {
        float s = 0.f;
        if (L.D)
            for (int b = 0; b < M; ++b) s += L.D[i * MM + a * M + b] * xi[b];
        else
            s = xi[a];
        r[a] = fl[i * M + a] - s;
    }
    int idx[3];
    cell_coords(L.g, i, idx);
    for (int d = 0; d < 2 * L.g.dim; ++d) {
        int64_t j;
        if (!stencil_neighbor(L.g, idx, i, d, j)) continue;
        float xj[M];
        corrected(j, xj);
        for (int a = 0; a < M; ++a) {
            float s = 0.f;
            for (int b = 0; b < M; ++b) s += L.S[(int64_t(d) * MM + a * M + b) * L.g.N + i] * xj[b];
            r[a] -= s;
        }
    }
    float z[M];
    mg_jacobi<M>(L, i, r, z);
    for (int a = 0; a < M; ++a) out[i * M + a] = xi[a] + z[a];
}

// Coarsest level (at most kMgCoarseCells cells): kMgCoarseSweeps damped block-Jacobi sweeps from
// zero, by one thread; the result replaces L.f
template <int M>
DMX_HD void mg_coarsest(const MgLevel& L)
{
    const int64_t N = L.g.N;
    float x[kMgCoarseCells * M], xn[kMgCoarseCells * M];
    for (int64_t i = 0; i < N; ++i) {
        float xi[M];
        mg_presmooth_value<M>(L, L.f, i, xi);
        for (int a = 0; a < M; ++a) x[i * M + a] = xi[a];
    }
    for (int sweep = 1; sweep < kMgCoarseSweeps; ++sweep) {
        for (int64_t i = 0; i < N; ++i) {
            float y[M], r[M], z[M];
            mg_apply_row<M>(L, x, i, y);
            for (int a = 0; a < M; ++a) r[a] = L.f[i * M + a] - y[a];
            mg_jacobi<M>(L, i, r, z);
            for (int a = 0; a < M; ++a) xn[i * M + a] = x[i * M + a] + z[a];
        }
        for (int64_t k = 0; k < N * M; ++k) x[k] = xn[k];
    }
    for (int64_t k = 0; k < N * M; ++k) L.f[k] = x[k];
}

// one V-cycle y = M^-1 v on the levels lev[0..nl-1] (host reference of the device sequence:
// down-sweeps, coarsest, up-sweeps, each over all cells of its level)
template <int M>
inline void mg_vcycle_host(const MgLevel* lev, int nl, const float* v, float* y)
{
    for (int l = 0; l + 1 < nl; ++l) {
        const float* fl = l == 0 ? v : lev[l].f;
        for (int64_t C = 0; C < lev[l + 1].g.N; ++C) mg_down<M>(lev[l], fl, lev[l + 1], C);
    }
    mg_coarsest<M>(lev[nl - 1]);
    for (int l = nl - 2; l >= 0; --l) {
        const float* fl = l == 0 ? v : lev[l].f;
        float* out = l == 0 ? y : lev[l].f;
        for (int64_t i = 0; i < lev[l].g.N; ++i) mg_up<M>(lev[l], fl, lev[l + 1], lev[l + 1].f, i, out);
    }
}

} // namespace dmx
