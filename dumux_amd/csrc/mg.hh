// Aggregation multigrid preconditioner of the batched FP32 BiCGSTAB (host + device).
//
// The reference offers algebraic multigrid as a BiCGSTAB preconditioner
// (AMGBiCGSTABIstlSolver, linear/istlsolvers.hh:729-736; dune-istl's aggregation AMG,
// third-party). On the structured grid the aggregates are fixed boxes of 2x2x2 cells (2x2 in
// 2D, fewer at the grid's end), so every level is again a 2*dim+1 point stencil of m x m blocks
// and keeps the batch operator's layout:
//   S_l[(d*M*M + q)*N_l + i]  off-diagonal blocks (d = 2*axis + side), D_l, D_l^-1 per cell.
// The batch iterates on the left-scaled system D^-1 A (unit diagonal: level 0). The coarse
// operators are Galerkin products of the UNSCALED Jacobian with piecewise-constant prolongation
// P (unsmoothed aggregation), accumulated in FP64 and stored in FP32:
//   A_1 = P^T A P from the FP64 BCRS Jacobian, A_{l+1} = P^T A_l P, i.e.
//   D_{l+1}[C]   = sum over children c of (D_l[c] + the blocks of c's faces inside C)
//   S_{l+1,d}[C] = sum over the children on C's d-face of S_{l,d}[c]
// and level 0's residual is restricted by P^T D (D the Jacobian's diagonal blocks), so that
// P^T D (D^-1 A) P = A_1 is the Galerkin operator of the scaled system too. (Products of the
// row-scaled operator itself sum rows of very different scales; near a saturation front their
// diagonal blocks become indefinite and block Jacobi diverges on them.)
// A V-cycle: one damped block-Jacobi pre-smoothing step from zero, the residual restricted by
// summation over the children, the coarse correction prolongated by injection, one
// post-smoothing step; the coarsest level (at most kMgCoarseCells cells) gets kMgCoarseSweeps
// damped block-Jacobi sweeps. FP32 inside the FP32 inner solve, in a fixed order: every sum
// over the children runs dz, dy, dx, so the host form and the device kernels give the same bits.
//
// Level sizes stay below 2^31 cells (the BCRS pattern indexes blocks with int32), so cell
// coordinates are computed with 32-bit divisions.
#pragma once
#include "stencil.hh"

#include <algorithm>

namespace dmx {

constexpr int kMgMaxLevels = 12;
constexpr int kMgCoarseCells = 8;
constexpr int kMgCoarseSweeps = 8;
constexpr int kMgTailCells = 512;   // first level of the one-workgroup tail: at most this many cells
constexpr int kMgTailLdsCells = 128; // cells of the tail's coarser levels together (kept in LDS)
constexpr int kMgBoxCells = 256;    // cells per box of the box-tiled level kernels
constexpr float kMgOmega = 0.8f;   // damping of the block-Jacobi smoother (0.8: ~10 % fewer BiCGSTAB iterations than 0.7 at 64^3)

struct MgLevel {
    GridDesc g;       // dim, n, stride, N of the level (geometry fields unused)
    TileGeom box;     // boxes of even extents (whole aggregates), at most kMgBoxCells cells
    float* S;         // 2*dim*m*m*N off-diagonal blocks
    float* D;         // m*m*N diagonal blocks (null on level 0: unit diagonal)
    float* Dinv;      // m*m*N inverse diagonal blocks (null on level 0)
    float* Dr;        // level 0: the unscaled diagonal blocks (restriction weights); null above
    float* x;         // m*N pre-smoothed iterate
    float* f;         // m*N right-hand side (levels >= 1; the result of the up-sweep replaces it)
    uint16_t* S16;    // level 0 in the batch: a bfloat16 copy of S (unit-diagonal scaled, O(1)
                      // entries) that the sweeps read instead of S, block-contiguous:
                      // S16[(d*N + i)*M*M + q] (one load per block); null: S
};

// bfloat16 <-> FP32 (round to nearest even; the copy only feeds the preconditioner)
DMX_HD uint16_t mg_to_bf16(float v)
{
    const uint32_t u = __builtin_bit_cast(uint32_t, v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);   // NaN stays NaN
    return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

DMX_HD float mg_from_bf16(uint16_t h) { return __builtin_bit_cast(float, uint32_t(h) << 16); }

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

// first level of the one-workgroup tail: at most kMgTailCells cells, and the levels below it at
// most kMgTailLdsCells cells together
DMX_HD int mg_tail_level(const GridDesc& g, int nl)
{
    GridDesc c = g;
    for (int l = 0; l + 1 < nl; ++l) {
        if (c.N <= kMgTailCells) {
            int64_t below = 0;
            GridDesc k = c;
            for (int j = l + 1; j < nl; ++j) {
                k = mg_coarse_grid(k);
                below += k.N;
            }
            if (below <= kMgTailLdsCells) return l;
        }
        c = mg_coarse_grid(c);
    }
    return nl - 1;
}

// FP32 values of a coarse level's LDS copy in the tail: S (2*dim directions), D, D^-1, x, f
DMX_HD int64_t mg_tail_floats(const GridDesc& g, int m)
{
    return (int64_t(2 * g.dim + 2) * m * m + 2 * m) * g.N;
}

// FP32 values of the multigrid buffer of a grid: level 0's bfloat16 operator copy (2*dim*m*m*N
// halves) and restriction weights, then S, D, D^-1, x, f per coarse level
inline int64_t mg_floats(const GridDesc& g, int m)
{
    const int nl = mg_num_levels(g);
    if (nl < 2) return 0;
    GridDesc c = g;
    int64_t t = int64_t(g.dim) * m * m * g.N + int64_t(m) * m * g.N;
    for (int l = 1; l < nl; ++l) {
        c = mg_coarse_grid(c);
        t += (int64_t(2) * c.dim * m * m + 2 * m * m + 2 * m) * c.N;
    }
    return t;
}

// algorithmic bytes of the down- (up = false) or up-sweep of level l of a grid: every array the
// sweep reads once, every output written once (the present off-diagonal blocks, level 0's at
// s0 bytes per value, the vectors, level 0's restriction weights, a coarse level's D and D^-1)
inline double mg_sweep_bytes(const GridDesc& L, int l, int m, double s0, bool up)
{
    const double mm = double(m) * m, v = 4.0 * m, N = double(L.N);
    double offd = 0.0;
    for (int a = 0; a < L.dim; ++a) offd += 2.0 * (L.n[a] - 1) * (N / L.n[a]);
    const double S = (l == 0 ? s0 : 4.0) * mm * offd, dd = l == 0 ? 0.0 : 2 * 4 * mm * N;
    const double Nc = double(mg_coarse_grid(L).N);
    if (up) return S + v * N + dd + v * N + v * Nc + v * N;      // S f (D D^-1) x e -> out
    return S + v * N + (l == 0 ? 4 * mm * N : dd) + v * N + v * Nc;   // S f (Dr | D D^-1) -> x, f_{l+1}
}

// algorithmic bytes of one V-cycle (mg_sweep_bytes of every level, the coarsest's S, D, D^-1
// and f read and written)
inline double mg_vcycle_bytes(const GridDesc& g, int m, double s0 = 4.0)
{
    const int nl = mg_num_levels(g);
    if (nl < 2) return 0.0;
    double tot = 0.0;
    GridDesc L = g;
    for (int l = 0; l + 1 < nl; ++l) {
        tot += mg_sweep_bytes(L, l, m, s0, false) + mg_sweep_bytes(L, l, m, s0, true);
        L = mg_coarse_grid(L);
    }
    const double mm = double(m) * m, N = double(L.N);
    double offd = 0.0;
    for (int a = 0; a < L.dim; ++a) offd += 2.0 * (L.n[a] - 1) * (N / L.n[a]);
    return tot + 4 * mm * offd + 2 * 4 * mm * N + 2 * 4.0 * m * N;
}

// boxes of a level: even extents, so that no aggregate straddles two boxes, x the longest (the
// loads of a row coalesce), at most kMgBoxCells cells
inline TileGeom mg_boxes(const GridDesc& g)
{
    auto even = [](int n) { return n + (n & 1); };
    TileGeom t{};
    if (g.dim == 3) {
        t.b[0] = std::min(64, even(g.n[0]));
        t.b[1] = std::min(((kMgBoxCells / t.b[0]) / 2) & ~1, even(g.n[1]));
        t.b[2] = std::min((kMgBoxCells / (t.b[0] * t.b[1])) & ~1, even(g.n[2]));
    } else {
        t.b[0] = std::min(128, even(g.n[0]));
        t.b[1] = std::min((kMgBoxCells / t.b[0]) & ~1, even(g.n[1]));
        t.b[2] = 1;
    }
    t.ntiles = 1;
    for (int a = 0; a < 3; ++a) {
        t.nt[a] = (g.n[a] + t.b[a] - 1) / t.b[a];
        t.ntiles *= t.nt[a];
    }
    return t;
}

// cell index of coordinates on a level
DMX_HD int64_t mg_index(const GridDesc& g, int x, int y, int z)
{
    return int64_t(z) * g.stride[2] + int64_t(y) * g.stride[1] + x;
}

// coordinates of cell i of a level (32-bit divisions)
DMX_HD void mg_coords(const GridDesc& g, int64_t i, int (&idx)[3])
{
    const uint32_t n0 = uint32_t(g.n[0]), n1 = uint32_t(g.n[1]), k = uint32_t(i);
    const uint32_t q = k / n0;
    const uint32_t z = q / n1;
    idx[0] = int(k - q * n0);
    idx[1] = int(q - z * n1);
    idx[2] = int(z);
}

// cell of thread t of box `box` (coordinates; false: the thread has no cell of the grid)
DMX_HD bool mg_box_cell(const GridDesc& g, const TileGeom& b, int32_t box, int t, int (&idx)[3])
{
    if (t >= b.b[0] * b.b[1] * b.b[2]) return false;
    const int q = box / b.nt[0];
    const int bz = q / b.nt[1];
    const int r = t / b.b[0];
    const int lz = r / b.b[1];
    idx[0] = (box - q * b.nt[0]) * b.b[0] + (t - r * b.b[0]);
    idx[1] = (q - bz * b.nt[1]) * b.b[1] + (r - lz * b.b[1]);
    idx[2] = bz * b.b[2] + lz;
    return idx[0] < g.n[0] && idx[1] < g.n[1] && idx[2] < g.n[2];
}

// cell of thread t of box `box` in aggregate-major order: lanes t with the same t / 2^dim hold
// the children of one aggregate, child t mod 2^dim = dz*4 + dy*2 + dx (mg_down's order);
// false: no cell of the grid (also for threads past the box). cidx: the aggregate's coordinates.
DMX_HD bool mg_box_child(const GridDesc& g, const TileGeom& b, int32_t box, int t, int (&idx)[3], int (&cidx)[3])
{
    const int lg = g.dim == 3 ? 3 : 2;
    const int c = t & ((1 << lg) - 1), a = t >> lg;
    const int h0 = b.b[0] >> 1, h1 = b.b[1] >> 1, h2 = b.b[2] > 1 ? b.b[2] >> 1 : 1;
    const int q = box / b.nt[0], bz = q / b.nt[1];
    const int r = a / h0, az = r / h1;
    cidx[0] = (box - q * b.nt[0]) * h0 + (a - r * h0);
    cidx[1] = (q - bz * b.nt[1]) * h1 + (r - az * h1);
    cidx[2] = bz * h2 + az;
    idx[0] = 2 * cidx[0] + (c & 1);
    idx[1] = 2 * cidx[1] + ((c >> 1) & 1);
    idx[2] = 2 * cidx[2] + (c >> 2);
    return az < h2 && idx[0] < g.n[0] && idx[1] < g.n[1] && idx[2] < g.n[2];
}

// Coarse blocks of one aggregate, from sums in FP64: D, D^-1 (inverted in FP64) and S in FP32
template <int M>
DMX_HD void mg_store_coarse(const MgLevel& c, int64_t C, const double (&Dc)[M * M], const double (&Sc)[6][M * M], int nd)
{
    constexpr int MM = M * M;
    for (int q = 0; q < MM; ++q) c.D[C * MM + q] = float(Dc[q]);
    float* Di = c.Dinv + C * MM;
    if (M == 1) {
        Di[0] = float(1.0 / Dc[0]);
    } else {
        const double a = Dc[0], b = Dc[1 % MM], cc = Dc[2 % MM], d = Dc[MM - 1];
        const double inv = 1.0 / (a * d - b * cc);
        Di[0] = float(d * inv);
        Di[1 % MM] = float(-b * inv);
        Di[2 % MM] = float(-cc * inv);
        Di[MM - 1] = float(a * inv);
    }
    for (int d = 0; d < 6; ++d)
        if (d < nd)
            for (int q = 0; q < MM; ++q) c.S[(int64_t(d) * MM + q) * c.g.N + C] = float(Sc[d][q]);
}

// Level 1 at coarse cell C (coordinates cidx) from the FP64 BCRS Jacobian of the fine grid fg
// (row i's blocks in column order, stencil_pos); also writes level 0's restriction weights
// Dr = float(J_ii) of the children
template <int M>
DMX_HD void mg_galerkin_jacobian(const GridDesc& fg, const int32_t* rowptr, const double* J, float* Dr, const MgLevel& c,
                                 const int (&cidx)[3], int64_t C)
{
    constexpr int MM = M * M;
    const int nd = 2 * fg.dim;
    double Dc[MM], Sc[6][MM];
    for (int q = 0; q < MM; ++q) Dc[q] = 0.0;
    for (int d = 0; d < 6; ++d)
        for (int q = 0; q < MM; ++q) Sc[d][q] = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int idx[3] = {2 * cidx[0] + dx, 2 * cidx[1] + dy, 2 * cidx[2] + dz};
                if (idx[0] >= fg.n[0] || idx[1] >= fg.n[1] || idx[2] >= fg.n[2]) continue;
                const int64_t i = mg_index(fg, idx[0], idx[1], idx[2]);
                const int64_t row0 = rowptr[i];
                auto pos = [&](int ax, int dir) {
                    return fg.dim == 3 ? stencil_pos<3>(fg, idx, ax, dir) : stencil_pos<2>(fg, idx, ax, dir);
                };
                const double* B = J + (row0 + pos(0, 0)) * MM;
                for (int q = 0; q < MM; ++q) {
                    Dc[q] += B[q];
                    Dr[i * MM + q] = float(B[q]);
                }
                for (int d = 0; d < 6; ++d) {   // fixed trip count: Sc stays in registers
                    const int ax = d >> 1;
                    const bool up = d & 1;
                    if (d >= nd || (up ? idx[ax] >= fg.n[ax] - 1 : idx[ax] <= 0)) continue;
                    const double* O = J + (row0 + pos(ax, up ? 1 : -1)) * MM;
                    const bool inside = up ? (idx[ax] & 1) == 0 : (idx[ax] & 1) == 1;
                    for (int q = 0; q < MM; ++q) {
                        if (inside) Dc[q] += O[q];
                        else Sc[d][q] += O[q];
                    }
                }
            }
    mg_store_coarse<M>(c, C, Dc, Sc, nd);
}

// Level l+1 >= 2 at coarse cell C (coordinates cidx) from level l (explicit diagonal)
template <int M>
DMX_HD void mg_galerkin(const MgLevel& f, const MgLevel& c, const int (&cidx)[3], int64_t C)
{
    constexpr int MM = M * M;
    const int nd = 2 * f.g.dim;
    double Dc[MM], Sc[6][MM];
    for (int q = 0; q < MM; ++q) Dc[q] = 0.0;
    for (int d = 0; d < 6; ++d)
        for (int q = 0; q < MM; ++q) Sc[d][q] = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int idx[3] = {2 * cidx[0] + dx, 2 * cidx[1] + dy, 2 * cidx[2] + dz};
                if (idx[0] >= f.g.n[0] || idx[1] >= f.g.n[1] || idx[2] >= f.g.n[2]) continue;
                const int64_t i = mg_index(f.g, idx[0], idx[1], idx[2]);
                for (int q = 0; q < MM; ++q) Dc[q] += double(f.D[i * MM + q]);
                for (int d = 0; d < 6; ++d) {
                    const int ax = d >> 1;
                    const bool up = d & 1;
                    if (d >= nd || (up ? idx[ax] >= f.g.n[ax] - 1 : idx[ax] <= 0)) continue;
                    const bool inside = up ? (idx[ax] & 1) == 0 : (idx[ax] & 1) == 1;
                    for (int q = 0; q < MM; ++q) {
                        const double s = f.S[(int64_t(d) * MM + q) * f.g.N + i];
                        if (inside) Dc[q] += s;
                        else Sc[d][q] += s;
                    }
                }
            }
    mg_store_coarse<M>(c, C, Dc, Sc, nd);
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

// pre-smoothing value of cell i: x = omega D^-1 f (an FP64 f is rounded to FP32 on load)
template <int M, typename F>
DMX_HD void mg_presmooth_value(const MgLevel& L, const F* f, int64_t i, float (&x)[M])
{
    float v[M];
    for (int a = 0; a < M; ++a) v[a] = float(f[i * M + a]);
    mg_jacobi<M>(L, i, v, x);
}

// r = f_i - D_i x_i (unit diagonal on level 0)
template <int M, typename F>
DMX_HD void mg_diag_residual(const MgLevel& L, const F* fl, int64_t i, const float (&xi)[M], float (&r)[M])
{
    for (int a = 0; a < M; ++a) {
        float s = 0.f;
        if (L.D)
            for (int b = 0; b < M; ++b) s += L.D[i * M * M + a * M + b] * xi[b];
        else
            s = xi[a];
        r[a] = float(fl[i * M + a]) - s;
    }
}

struct alignas(8) MgBf16x4 {
    uint32_t lo, hi;
};

// block (d, i) of the bfloat16 copy (M = 2: one 8-byte load)
template <int M>
DMX_HD void mg_load_bf16_block(const uint16_t* S16, int64_t N, int d, int64_t i, float (&B)[M * M])
{
    const uint16_t* blk = S16 + (int64_t(d) * N + i) * (M * M);
    if constexpr (M == 2) {
        const MgBf16x4 v = *reinterpret_cast<const MgBf16x4*>(blk);
        B[0] = __builtin_bit_cast(float, v.lo << 16);
        B[1] = __builtin_bit_cast(float, v.lo & 0xffff0000u);
        B[2] = __builtin_bit_cast(float, v.hi << 16);
        B[3] = __builtin_bit_cast(float, v.hi & 0xffff0000u);
    } else {
        for (int q = 0; q < M * M; ++q) B[q] = mg_from_bf16(blk[q]);
    }
}

// r -= S(d, i) xj
template <int M>
DMX_HD void mg_sub_offdiag(const MgLevel& L, int d, int64_t i, const float (&xj)[M], float (&r)[M])
{
    float B[M * M];
    if (L.S16)
        mg_load_bf16_block<M>(L.S16, L.g.N, d, i, B);
    else
        for (int q = 0; q < M * M; ++q) B[q] = L.S[(int64_t(d) * M * M + q) * L.g.N + i];
    for (int a = 0; a < M; ++a) {
        float s = 0.f;
        for (int b = 0; b < M; ++b) s += B[a * M + b] * xj[b];
        r[a] -= s;
    }
}

// Fine cell i (coordinates idx) in the down-sweep: the pre-smoothed x_i = omega D^-1 f_i (kept
// in L.x) and the residual r_i = f_i - (A x)_i, the neighbours' x recomputed from f (the first
// smoothing step is pointwise); level 0 returns D_i r_i (restriction by P^T D)
template <int M, typename F>
DMX_HD void mg_fine_residual(const MgLevel& L, const F* fl, const int (&idx)[3], int64_t i, float (&r)[M])
{
    float xi[M];
    mg_presmooth_value<M>(L, fl, i, xi);
    for (int a = 0; a < M; ++a) L.x[i * M + a] = xi[a];
    mg_diag_residual<M>(L, fl, i, xi, r);
    for (int d = 0; d < 6; ++d) {   // fixed trip count: the level's fields stay in registers
        int64_t j;
        if (d >= 2 * L.g.dim || !stencil_neighbor(L.g, idx, i, d, j)) continue;
        float xj[M];
        mg_presmooth_value<M>(L, fl, j, xj);
        mg_sub_offdiag<M>(L, d, i, xj, r);
    }
    if (L.Dr) {
        float t[M];
        for (int a = 0; a < M; ++a) {
            float s = 0.f;
            for (int b = 0; b < M; ++b) s += L.Dr[i * M * M + a * M + b] * r[b];
            t[a] = s;
        }
        for (int a = 0; a < M; ++a) r[a] = t[a];
    }
}

// Down-sweep at coarse cell C (coordinates cidx) of level l+1: f_{l+1}[C] = the children's
// residuals summed in the order dz, dy, dx
template <int M>
DMX_HD void mg_down(const MgLevel& L, const float* fl, const MgLevel& Cl, const int (&cidx)[3], int64_t C)
{
    float acc[M];
    for (int a = 0; a < M; ++a) acc[a] = 0.f;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int idx[3] = {2 * cidx[0] + dx, 2 * cidx[1] + dy, 2 * cidx[2] + dz};
                if (idx[0] >= L.g.n[0] || idx[1] >= L.g.n[1] || idx[2] >= L.g.n[2]) continue;
                float r[M];
                mg_fine_residual<M>(L, fl, idx, mg_index(L.g, idx[0], idx[1], idx[2]), r);
                for (int a = 0; a < M; ++a) acc[a] += r[a];
            }
    for (int a = 0; a < M; ++a) Cl.f[C * M + a] = acc[a];
}

// Up-sweep at fine cell i (coordinates idx) of level l: x = x_l + e[parent] (e = the coarse
// level's result on grid cg), then out[i] = x_i + omega D^-1 (f_l - A_l x)_i with the corrected
// neighbours
template <int M, typename F, typename O>
DMX_HD void mg_up(const MgLevel& L, const F* fl, const GridDesc& cg, const float* e, const int (&idx)[3], int64_t i, O* out)
{
    auto corrected = [&](const int (&c)[3], int64_t k, float (&v)[M]) {
        const int64_t P = mg_index(cg, c[0] >> 1, c[1] >> 1, c[2] >> 1);
        for (int a = 0; a < M; ++a) v[a] = L.x[k * M + a] + e[P * M + a];
    };
    float xi[M];
    corrected(idx, i, xi);
    float r[M];
    mg_diag_residual<M>(L, fl, i, xi, r);
    for (int d = 0; d < 6; ++d) {   // fixed trip count: the level's fields stay in registers
        int64_t j;
        if (d >= 2 * L.g.dim || !stencil_neighbor(L.g, idx, i, d, j)) continue;
        int jdx[3] = {idx[0], idx[1], idx[2]};
        jdx[d >> 1] += (d & 1) ? 1 : -1;
        float xj[M];
        corrected(jdx, j, xj);
        mg_sub_offdiag<M>(L, d, i, xj, r);
    }
    float z[M];
    mg_jacobi<M>(L, i, r, z);
    for (int a = 0; a < M; ++a) out[i * M + a] = O(xi[a] + z[a]);
}

// one coarsest-level sweep at cell i: xn = x_i + omega D^-1 (f - A x)_i (x: the whole level)
template <int M>
DMX_HD void mg_coarse_update(const MgLevel& L, const float* x, int64_t i, float (&xn)[M])
{
    float xi[M];
    for (int a = 0; a < M; ++a) xi[a] = x[i * M + a];
    float r[M];
    mg_diag_residual<M>(L, L.f, i, xi, r);
    int idx[3];
    mg_coords(L.g, i, idx);
    for (int d = 0; d < 6; ++d) {   // fixed trip count: the level's fields stay in registers
        int64_t j;
        if (d >= 2 * L.g.dim || !stencil_neighbor(L.g, idx, i, d, j)) continue;
        float xj[M];
        for (int b = 0; b < M; ++b) xj[b] = x[j * M + b];
        mg_sub_offdiag<M>(L, d, i, xj, r);
    }
    float z[M];
    mg_jacobi<M>(L, i, r, z);
    for (int a = 0; a < M; ++a) xn[a] = xi[a] + z[a];
}

// Coarsest level (at most kMgCoarseCells cells): kMgCoarseSweeps damped block-Jacobi sweeps from
// zero (the first is the pre-smoothing value), by one thread; the result replaces L.f
template <int M>
DMX_HD void mg_coarsest_host(const MgLevel& L)
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
            float v[M];
            mg_coarse_update<M>(L, x, i, v);
            for (int a = 0; a < M; ++a) xn[i * M + a] = v[a];
        }
        for (int64_t k = 0; k < N * M; ++k) x[k] = xn[k];
    }
    for (int64_t k = 0; k < N * M; ++k) L.f[k] = x[k];
}

// one V-cycle y = M^-1 v on the levels lev[0..nl-1] (host form of the device sequence)
template <int M>
inline void mg_vcycle_host(const MgLevel* lev, int nl, const float* v, float* y)
{
    for (int l = 0; l + 1 < nl; ++l) {
        const float* fl = l == 0 ? v : lev[l].f;
        for (int64_t C = 0; C < lev[l + 1].g.N; ++C) {
            int cidx[3];
            mg_coords(lev[l + 1].g, C, cidx);
            mg_down<M>(lev[l], fl, lev[l + 1], cidx, C);
        }
    }
    mg_coarsest_host<M>(lev[nl - 1]);
    for (int l = nl - 2; l >= 0; --l) {
        const float* fl = l == 0 ? v : lev[l].f;
        float* out = l == 0 ? y : lev[l].f;
        for (int64_t i = 0; i < lev[l].g.N; ++i) {
            int idx[3];
            mg_coords(lev[l].g, i, idx);
            mg_up<M>(lev[l], fl, lev[l + 1].g, lev[l + 1].f, idx, i, out);
        }
    }
}

} // namespace dmx
