// Host-side problem setup: grid and physics descriptors, BCRS pattern, ILU(0) level
// schedule and the seeded synthetic problems of the benchmark workloads.
#pragma once
#include "cellops.hh"
#include "../../include/dumux_amd.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace dmx {

inline GridDesc make_grid(const dmx_problem_desc& d)
{
    GridDesc g{};
    g.dim = d.dim;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = a < d.dim ? d.n[a] : 1;
        g.lower[a] = d.lower[a];
        // YaspGrid EquidistantOffsetCoordinates: h = (upper - lower) / n
        g.h[a] = a < d.dim ? (d.upper[a] - d.lower[a]) / g.n[a] : 0.0;
    }
    g.N = int64_t(g.n[0]) * g.n[1] * g.n[2];
    g.stride[0] = 1;
    g.stride[1] = g.n[0];
    g.stride[2] = int64_t(g.n[0]) * g.n[1];
    int64_t off = 0;
    for (int s = 0; s < 7; ++s) g.sideOffset[s] = 0;
    for (int s = 0; s < 2 * g.dim; ++s) {
        g.sideOffset[s] = off;
        off += g.N / g.n[s / 2];
    }
    g.sideOffset[2 * g.dim] = off;
    return g;
}

inline int64_t num_boundary_faces(const GridDesc& g) { return g.sideOffset[2 * g.dim]; }

inline Phys make_phys(const dmx_problem_desc& d)
{
    Phys p{};
    p.m = d.model == DMX_MODEL_TWOP ? 2 : 1;
    p.gravity = d.gravity;
    for (int a = 0; a < 2; ++a) { p.rho[a] = d.rho[a]; p.mu[a] = d.mu[a]; }
    p.phi = d.porosity;
    p.fdBase = d.fd_base_eps;
    p.fdMag[0] = d.fd_magnitude[0];
    p.fdMag[1] = d.fd_magnitude[1];
    if (p.m == 2)
    {
        vg_init(p.law[0], d.vg_alpha, d.vg_n, d.vg_l, d.swr, d.snr, d.pc_low_swe, d.pc_high_swe,
                d.krn_low_swe, d.krw_high_swe);
        const dmx_vg_material& q = d.material1;
        if (d.material_id)
            vg_init(p.law[1], q.vg_alpha, q.vg_n, q.vg_l, q.swr, q.snr, q.pc_low_swe, q.pc_high_swe, q.krn_low_swe,
                    q.krw_high_swe);
        else
            p.law[1] = p.law[0];
    }
    return p;
}

// getJacobianPattern<true> (assembly/jacobianpattern.hh:25-53) for the 2*dim+1 point TPFA
// stencil: row I holds I and its face neighbours, column indices ascending.
inline void build_pattern(const GridDesc& g, std::vector<int32_t>& rowptr, std::vector<int32_t>& colidx)
{
    rowptr.assign(g.N + 1, 0);
    colidx.clear();
    colidx.reserve(size_t(g.N) * (2 * g.dim + 1));
    int64_t row[7];
    for (int64_t I = 0; I < g.N; ++I) {
        int idx[3] = {int(I % g.n[0]), int((I / g.n[0]) % g.n[1]), int(I / (int64_t(g.n[0]) * g.n[1]))};
        int c = 0;
        row[c++] = I;
        for (int a = 0; a < g.dim; ++a) {
            if (idx[a] > 0) row[c++] = I - g.stride[a];
            if (idx[a] < g.n[a] - 1) row[c++] = I + g.stride[a];
        }
        std::sort(row, row + c);
        for (int k = 0; k < c; ++k) colidx.push_back(int32_t(row[k]));
        rowptr[I + 1] = int32_t(colidx.size());
    }
}

// Level schedule of the lower-triangular dependency graph: level(i) = 1 + max level of the
// columns k < i of row i (0 without any). Rows of one level are independent in the ILU(0)
// factorization and in both triangular sweeps; rows of a level are listed ascending.
inline int level_schedule(int64_t N, const int32_t* rowptr, const int32_t* colidx,
                          std::vector<int32_t>& level_ptr, std::vector<int32_t>& level_rows)
{
    std::vector<int32_t> level(N, 0);
    int32_t nlev = 0;
    for (int64_t i = 0; i < N; ++i) {
        int32_t l = 0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1] && colidx[k] < i; ++k)
            l = std::max(l, level[colidx[k]] + 1);
        level[i] = l;
        nlev = std::max(nlev, l + 1);
    }
    level_ptr.assign(nlev + 1, 0);
    for (int64_t i = 0; i < N; ++i) level_ptr[level[i] + 1]++;
    for (int32_t l = 0; l < nlev; ++l) level_ptr[l + 1] += level_ptr[l];
    std::vector<int32_t> fill(level_ptr.begin(), level_ptr.end() - 1);
    level_rows.assign(N, 0);
    for (int64_t i = 0; i < N; ++i) level_rows[fill[level[i]]++] = int32_t(i);
    return nlev;
}

// Material of every boundary face: that of its inside cell (material_id may be null: all 0)
inline void boundary_materials(const GridDesc& g, const int32_t* material_id, std::vector<int32_t>& bmat)
{
    bmat.assign(size_t(num_boundary_faces(g)), 0);
    if (!material_id) return;
    for (int64_t i = 0; i < g.N; ++i) {
        const int idx[3] = {int(i % g.n[0]), int((i / g.n[0]) % g.n[1]), int(i / (int64_t(g.n[0]) * g.n[1]))};
        for (int f = 0; f < 2 * g.dim; ++f) {
            const int ax = f / 2;
            const bool boundary = (f & 1) ? idx[ax] == g.n[ax] - 1 : idx[ax] == 0;
            if (boundary) bmat[size_t(boundary_face(g, idx, f))] = material_id[i];
        }
    }
}

// True if row i holds exactly the cell itself and its face neighbours on g (ascending), the
// pattern the box sweeps of the ILU(0) apply rely on.
inline bool structured_stencil_pattern(const GridDesc& g, const std::vector<int32_t>& rowptr,
                                       const std::vector<int32_t>& colidx)
{
    std::vector<int64_t> want;
    for (int64_t i = 0; i < g.N; ++i) {
        const int64_t c[3] = {i % g.n[0], (i / g.n[0]) % g.n[1], i / (int64_t(g.n[0]) * g.n[1])};
        want.clear();
        for (int a = 2; a >= 0; --a)
            if (c[a] > 0) want.push_back(i - g.stride[a]);
        want.push_back(i);
        for (int a = 0; a < 3; ++a)
            if (c[a] + 1 < g.n[a]) want.push_back(i + g.stride[a]);
        if (int64_t(rowptr[i + 1] - rowptr[i]) != int64_t(want.size())) return false;
        for (size_t k = 0; k < want.size(); ++k)
            if (colidx[rowptr[i] + k] != want[k]) return false;
    }
    return true;
}

// Boxes for the ILU(0) sweeps and their order: ascending box level bx+by+bz, then box index. A
// box depends only on boxes of lower level (forward sweep) or higher level (backward sweep).
// A sweep costs about (box levels) x (levels inside a box) barrier steps plus one wait between
// workgroups per box level; measured on 64^3 (2x2 blocks) 4x4x4 boxes are fastest (8x8x8:
// 1.45x slower), on 100^2 (1x1) 16x16 boxes (8x8: 1.06x slower). Boxes are clipped at the
// grid's end.
inline TileGeom sweep_boxes(const GridDesc& g, std::vector<int32_t>& order)
{
    TileGeom t{};
    const int b3[3] = {4, 4, 4}, b2[3] = {16, 16, 1};
    for (int a = 0; a < 3; ++a) {
        t.b[a] = std::min(g.dim == 3 ? b3[a] : b2[a], g.n[a]);
        t.nt[a] = (g.n[a] + t.b[a] - 1) / t.b[a];
    }
    t.ntiles = t.nt[0] * t.nt[1] * t.nt[2];
    order.resize(t.ntiles);
    for (int32_t k = 0; k < t.ntiles; ++k) order[k] = k;
    auto level = [&](int32_t k) { return k % t.nt[0] + (k / t.nt[0]) % t.nt[1] + k / (t.nt[0] * t.nt[1]); };
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return level(x) < level(y); });
    return t;
}

// ---- seeded permeability ----------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) keyed (seed, problem id), counter (cell, attempt,
// 0x4B, 0); two 53-bit uniforms in (-1,1) feed Marsaglia's polar method; K = 10^(mu + sigma z).
inline void philox4x32_10(const uint32_t in[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t x0 = in[0], x1 = in[1], x2 = in[2], x3 = in[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t a = uint64_t(0xD2511F53u) * x0;
        const uint64_t b = uint64_t(0xCD9E8D57u) * x2;
        const uint32_t y0 = uint32_t(b >> 32) ^ x1 ^ k0;
        const uint32_t y2 = uint32_t(a >> 32) ^ x3 ^ k1;
        x1 = uint32_t(b);
        x3 = uint32_t(a);
        x0 = y0;
        x2 = y2;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

inline double lognormal_permeability(uint32_t seed, uint32_t pid, uint32_t cell, double mu, double sigma)
{
    const uint32_t key[2] = {seed, pid};
    for (uint32_t attempt = 0; attempt < 1000; ++attempt) {
        const uint32_t ctr[4] = {cell, attempt, 0x4Bu, 0u};
        uint32_t w[4];
        philox4x32_10(ctr, key, w);
        const double u1 = double((((uint64_t)w[0] << 32) | w[1]) >> 11);
        const double u2 = double((((uint64_t)w[2] << 32) | w[3]) >> 11);
        const double v1 = (u1 + 0.5) * 0x1.0p-52 - 1.0;
        const double v2 = (u2 + 0.5) * 0x1.0p-52 - 1.0;
        const double s = v1 * v1 + v2 * v2;
        if (s >= 1.0 || s == 0.0) continue;
        const double z = v1 * std::sqrt(-2.0 * dmx::log(s) / s);
        return dmx::pow(10.0, mu + sigma * z);
    }
    return dmx::pow(10.0, mu);
}

// Synthetic problems; boundary/initial values mirror
// test/porousmediumflow/2p/incompressible/problem.hh:55-122 (hydrostatic pw = p_ref - rho_w*g*depth).
inline int build_problem(const dmx_build_spec& S, dmx_problem_desc& P, double* K, int32_t* bc_type,
                         double* bc_value, double* u0)
{
    if (S.dim < 2 || S.dim > 3 || (S.kind != 1 && S.kind != 2)) return DMX_ERR_INVALID;
    std::memset(&P, 0, sizeof(P));
    P.dim = S.dim;
    P.model = S.kind == 2 ? DMX_MODEL_TWOP : DMX_MODEL_ONEP;
    for (int a = 0; a < 3; ++a) {
        P.n[a] = a < S.dim ? S.n[a] : 1;
        if (P.n[a] < 1) return DMX_ERR_INVALID;
        P.lower[a] = 0.0;
        P.upper[a] = a < S.dim ? 1.0 : 0.0;
    }
    P.gravity = S.gravity;
    P.rho[0] = 1000.0;  P.mu[0] = 1e-3;   // SimpleH2O (material/components/simpleh2o.hh:265-312)
    P.rho[1] = 1460.0;  P.mu[1] = 5.7e-4; // Trichloroethene (material/components/trichloroethene.hh:148-172)
    P.porosity = 0.2;
    P.vg_alpha = 4e-4; P.vg_n = 2.5; P.vg_l = 0.5; P.swr = 0.05; P.snr = 0.0;
    P.pc_low_swe = 0.01; P.pc_high_swe = 0.99; P.krn_low_swe = 0.1; P.krw_high_swe = 0.9;
    P.fd_base_eps = 1e-10;
    P.fd_magnitude[0] = -1.0;
    P.fd_magnitude[1] = -1.0;
    P.diff_method = DMX_DIFF_NUMERIC;
    P.material_id = nullptr;
    P.material1 = dmx_vg_material{P.vg_alpha, P.vg_n, P.vg_l, P.swr, P.snr,
                                  P.pc_low_swe, P.pc_high_swe, P.krn_low_swe, P.krw_high_swe};
    const GridDesc g = make_grid(P);
    const int m = P.model == DMX_MODEL_TWOP ? 2 : 1;
    // counter-based generator: cells are independent, so large fields are filled by several
    // host threads with identical results
    auto fill = [&](int64_t c0, int64_t c1) {
        for (int64_t c = c0; c < c1; ++c)
            K[c] = lognormal_permeability(S.seed, S.pid, uint32_t(c), S.mu_log10K, S.sigma_log10K);
    };
    const int64_t nthreads = g.N < (int64_t(1) << 16) ? 1 : std::min<int64_t>(8, std::max(1u, std::thread::hardware_concurrency()));
    if (nthreads == 1) {
        fill(0, g.N);
    } else {
        std::vector<std::thread> pool;
        const int64_t chunk = (g.N + nthreads - 1) / nthreads;
        for (int64_t t = 0; t < nthreads; ++t)
            pool.emplace_back(fill, t * chunk, std::min(g.N, (t + 1) * chunk));
        for (auto& th : pool) th.join();
    }
    const int vert = g.dim - 1;
    const double gv = S.gravity ? -9.81 : 0.0;
    const double top = P.upper[vert];
    for (int f = 0; f < 2 * g.dim; ++f) {
        const int ax = f / 2, side = f % 2;
        int t[2] = {-1, -1}, nt = 0;
        for (int a = 0; a < g.dim; ++a)
            if (a != ax) t[nt++] = a;
        const int64_t count = g.N / g.n[ax];
        for (int64_t local = 0; local < count; ++local) {
            int idx[3] = {0, 0, 0};
            idx[t[0]] = int(local % g.n[t[0]]);
            if (nt > 1) idx[t[1]] = int(local / g.n[t[0]]);
            idx[ax] = side ? g.n[ax] - 1 : 0;
            const int64_t b = g.sideOffset[f] + local;
            // face centre
            double ip[3] = {0, 0, 0};
            for (int a = 0; a < g.dim; ++a) {
                const double lo = gcoord(g, a, idx[a]);
                const double up = gcoord(g, a, idx[a] + 1);
                ip[a] = a == ax ? (side ? up : lo) : lo + 0.5 * (up - lo);
            }
            double* v = bc_value + b * m;
            if (S.kind == 1) {
                if (ax == 0) { bc_type[b] = 1; v[0] = side ? S.p_right : S.p_left; }
                else { bc_type[b] = 0; v[0] = 0.0; }
            } else if (ax < vert) {
                bc_type[b] = 1;
                const double depth = top - ip[vert];
                v[0] = S.p_ref - P.rho[0] * gv * depth;
                v[1] = 0.0;
            } else {
                bc_type[b] = 0;
                v[0] = 0.0;
                v[1] = 0.0;
                if (side == 1) {
                    bool inlet = true;
                    for (int a = 0; a < g.dim; ++a) {
                        if (a == ax) continue;
                        const double L = P.upper[a] - P.lower[a];
                        if (!(P.lower[a] + S.inlet_lo * L < ip[a] && ip[a] < P.lower[a] + S.inlet_hi * L)) inlet = false;
                    }
                    if (inlet) v[1] = S.influx;
                }
            }
        }
    }
    for (int64_t c = 0; c < g.N; ++c) {
        if (S.kind == 1) {
            u0[c] = S.p_ref;
        } else {
            const int kv = int(vert == 2 ? c / (int64_t(g.n[0]) * g.n[1]) : (c / g.n[0]) % g.n[1]);
            const double lo = gcoord(g, vert, kv);
            const double up = gcoord(g, vert, kv + 1);
            const double depth = top - (lo + 0.5 * (up - lo));
            u0[2 * c] = S.p_ref - P.rho[0] * gv * depth;
            u0[2 * c + 1] = 0;
        }
    }
    P.K = K;
    P.bc_type = bc_type;
    P.bc_value = bc_value;
    return DMX_OK;
}

} // namespace dmx
