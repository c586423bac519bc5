// Per-cell CCTpfa residual and forward-difference Jacobian column (host + device).
//
// One call handles one cell I ("column owner"): it evaluates the cell residual R(I) and, for
// each primary variable of I, the deflected residual of I and the deflected fluxes of every
// face neighbour J across the face shared with I. It writes r[I], the diagonal block A[I][I]
// and the off-diagonal blocks A[J][I]; no other cell writes these entries, so a grid of
// threads needs no atomics.
//
// Reference arithmetic restated here (operation order preserved):
//   CCLocalAssembler<numeric>::assembleJacobianAndResidualImpl  assembly/cclocalassembler.hh:164-348
//   NumericDifferentiation::partialDerivative (forward)        common/numericdifferentiation.hh:66-123
//   NumericEpsilon (baseEps 1e-10, magnitude from |x|+1)       assembly/numericepsilon.hh:37-53
//   FVLocalResidual::evalFluxAndSource / evalStorage            assembly/fvlocalresidual.hh:150-169, 274-333
//   CCLocalResidual::evalFlux (interior/Dirichlet/Neumann)      assembly/cclocalresidual.hh:49-105
//   CCTpfaDarcysLaw::flux / calculateTransmissibility           flux/cctpfa/darcyslaw.hh:153-259
//   computeTpfaTransmissibility (scalar K)                      discretization/cellcentered/tpfa/computetransmissibility.hh:66-80
//   upwind with Flux.UpwindWeight = 1                           flux/upwindscheme.hh:35-54, 113-122
//   gravity g = -9.81 e_dim                                     common/fvspatialparams.hh:48-51
// Geometry (YaspGrid convention): x_a(i) = lower_a + i*h_a, centre = lo + 0.5*(up - lo), face
// centre = cell centre with the normal coordinate replaced. With axis-aligned unit normals the
// reference's vector loops reduce exactly (additions of +0) to the scalar forms used below.
#pragma once
#include "physics.hh"

namespace dmx {

struct GridDesc {
    int dim;
    int n[3];
    double lower[3];
    double h[3];
    int64_t N;
    int64_t stride[3];
    int64_t sideOffset[7];
};

DMX_HD double gcoord(const GridDesc& g, int a, int i) { return g.lower[a] + i * g.h[a]; }

// Boxes of the structured grid for the one-launch ILU(0) sweeps: extent b per axis (at most
// kTileCells cells, (b0+1)(b1+1)(b2+1) <= kTileLds), nt boxes per axis.
constexpr int kTileCells = 512;
constexpr int kTileLds = 729;
constexpr int kIluSweepMaxSide = 3;   // off-diagonal blocks per triangle of a row (7-point stencil)
struct TileGeom {
    int b[3];
    int nt[3];
    int32_t ntiles;
};

// Arrays of one problem (all device pointers in kernels).
struct AsmPtrs {
    const double* u;         // N*m, current iterate
    const double* uprev;     // N*m, previous time level (null: stationary)
    const double* K;         // N
    const int32_t* bc_type;  // per boundary face: 0 Neumann, 1 Dirichlet
    const double* bc_value;  // per boundary face * m
    const double* ghost_p;   // per boundary face * m (Dirichlet ghost phase pressures)
    const double* ghost_up;  // per boundary face * m (Dirichlet ghost rho*lambda)
    const double* vv_p;      // N*m undeflected phase pressures
    const double* vv_up;     // N*m undeflected rho*lambda
    const double* vv_d;      // N*3 (2p analytic Jacobian): -dkrw/dsw, -dkrn/dsw, -dpc/dsw at Sw
    const int32_t* mat;      // N material ids (null: all material 0)
    const int32_t* rowptr;   // BCRS row pointer (block rows)
    double* r;               // N*m residual
    double* J;               // nnzb*m*m Jacobian values
    float* S32;              // optional FP32 off-diagonal blocks, structured-stencil layout
                             // S32[(d*m*m + q)*N + row], d = direction from the row to the column
    float* Dinv32;           // optional FP32 inverse of the diagonal block (block-Jacobi)
    double* Dinv64;          // optional FP64 inverse of the diagonal block
    int32_t* fail;           // set to 1 if a diagonal block is singular (with Dinv32)
    double dt;
    int64_t rowBase;         // global block-row offset of this problem (batched BCRS)
    int analytic;            // 1: hand-coded derivatives (CCLocalAssembler<analytic>), 0: forward differences
    const uint8_t* green;    // partial reassembly: 1 keeps cell I's Jacobian column (residual only);
                             // null: every column is assembled
};

// boundary face index of the face on side f of the cell with coordinates idx
DMX_HD int64_t boundary_face(const GridDesc& g, const int idx[3], int f)
{
    const int ax = f / 2;
    int t0 = -1, t1 = -1;
    for (int a = 0; a < g.dim; ++a) {
        if (a == ax) continue;
        if (t0 < 0) t0 = a; else t1 = a;
    }
    int64_t local = idx[t0];
    if (t1 >= 0) local += int64_t(idx[t1]) * g.n[t0];
    return g.sideOffset[f] + local;
}

// Position of column (row J +/- stride_ax) inside the sorted row J of the 2*dim+1 point stencil.
// dir = -1: the column is the lower neighbour of J along ax; +1: upper; 0: the diagonal.
template <int DIM>
DMX_HD int stencil_pos(const GridDesc& g, const int jdx[3], int ax, int dir)
{
    int lower = 0;
    for (int a = 0; a < DIM; ++a) lower += jdx[a] > 0;
    if (dir == 0) return lower;
    if (dir < 0) {
        int pos = 0;
        for (int a = DIM - 1; a > ax; --a) pos += jdx[a] > 0;
        return pos;
    }
    int pos = lower + 1;
    for (int a = 0; a < ax; ++a) pos += jdx[a] < g.n[a] - 1;
    return pos;
}

struct FaceData {
    int kind;       // 0 interior, 1 Dirichlet, 2 Neumann
    double area;
    double tij;
    double ti, tj;  // inside / outside transmissibility (tj = outsideTi)
    double aI, aJ;  // K*(n.g) of I and of the outside cell (I's normal)
    double bI, bJ;  // same with the neighbour's normal (-n)
    CellState o;    // outside state (neighbour or Dirichlet ghost)
    double gN[2];   // Neumann flux per area
    int64_t slot;   // block slot A[J][I]
};

template <int M>
DMX_HD double upwind_flux(double F, double upIn, double upOut)
{
    const double w = 1.0;
    if (__builtin_signbit(F)) return F * (w * upOut + (1.0 - w) * upIn);
    return F * (w * upIn + (1.0 - w) * upOut);
}

// Darcy flux of phase a across interior or Dirichlet face fd seen from I (before upwinding),
// darcyslaw.hh:153-213
DMX_HD double darcy_flux_I(const Phys& ph, const FaceData& fd, const CellState& sI, int a)
{
    double F;
    if (ph.gravity) {
        if (fd.kind == 0) {
            const double rho = (ph.rho[a] + ph.rho[a]) * 0.5;
            F = fd.tij * (sI.p[a] - fd.o.p[a]) + rho * fd.area * fd.aI;
            F -= rho * fd.tij / fd.tj * (fd.aI - fd.aJ);
        } else {
            const double rho = ph.rho[a];
            F = fd.tij * (sI.p[a] - fd.o.p[a]) + rho * fd.area * fd.aI;
        }
    } else {
        F = fd.tij * (sI.p[a] - fd.o.p[a]);
    }
    return F;
}

// flux vector of face fd seen from I (inside) with I's state sI
template <int M>
DMX_HD void face_flux_I(const Phys& ph, const FaceData& fd, const CellState& sI, double* out)
{
    for (int a = 0; a < M; ++a) {
        double fl;
        if (fd.kind == 2) fl = fd.gN[a] * (fd.area * 1.0);
        else fl = upwind_flux<M>(darcy_flux_I(ph, fd, sI, a), sI.up[a], fd.o.up[a]);
        out[a] = 0.0 + fl;
    }
}

// Derivatives of the 2p fluxes across face fd of I w.r.t. (pw, Sn) of I (dII) and of the
// neighbour (dIJ, interior faces): TwoPIncompressibleLocalResidual::addFluxDerivatives
// (porousmediumflow/2p/incompressiblelocalresidual.hh:156-246) and addCCDirichletFluxDerivatives
// (:420-485), Flux.UpwindWeight = 1. dI / dO: -dkrw/dsw, -dkrn/dsw, -dpc/dsw of I / the
// neighbour at their wetting saturations.
DMX_HD void twop_flux_derivatives(const Phys& ph, const FaceData& fd, const CellState& sI, const double* dI,
                                  const double* dO, double* dII, double* dIJ)
{
    const double upwindWeight = 1.0;
    const double flux_w = darcy_flux_I(ph, fd, sI, 0);
    const double flux_n = darcy_flux_I(ph, fd, sI, 1);
    const double insideWeight_w = __builtin_signbit(flux_w) ? (1.0 - upwindWeight) : upwindWeight;
    const double outsideWeight_w = 1.0 - insideWeight_w;
    const double insideWeight_n = __builtin_signbit(flux_n) ? (1.0 - upwindWeight) : upwindWeight;
    const double outsideWeight_n = 1.0 - insideWeight_n;
    const double rhow_muw = ph.rho[0] / ph.mu[0];
    const double rhon_mun = ph.rho[1] / ph.mu[1];
    // rho * mobility of I and of the outside cell (or Dirichlet ghost)
    const double rhowKrw_muw_inside = sI.up[0], rhonKrn_mun_inside = sI.up[1];
    const double rhowKrw_muw_outside = fd.o.up[0], rhonKrn_mun_outside = fd.o.up[1];
    const double up_w = rhowKrw_muw_inside * insideWeight_w + rhowKrw_muw_outside * outsideWeight_w;
    const double up_n = rhonKrn_mun_inside * insideWeight_n + rhonKrn_mun_outside * outsideWeight_n;
    if (fd.kind == 0) {
        const double rho_mu_flux_w = rhow_muw * flux_w;
        const double rho_mu_flux_n = rhon_mun * flux_n;
        const double tij_up_w = fd.tij * up_w;
        const double tij_up_n = fd.tij * up_n;
        dII[0] += tij_up_w;
        dIJ[0] -= tij_up_w;
        dII[1] += rho_mu_flux_w * dI[0] * insideWeight_w;
        dIJ[1] += rho_mu_flux_w * dO[0] * outsideWeight_w;
        dII[2] += tij_up_n;
        dIJ[2] -= tij_up_n;
        dII[3] += rho_mu_flux_n * dI[1] * insideWeight_n;
        dIJ[3] += rho_mu_flux_n * dO[1] * outsideWeight_n;
        dII[3] += tij_up_n * dI[2];
        dIJ[3] -= tij_up_n * dO[2];
    } else {
        dII[0] += fd.tij * up_w;
        dII[1] += rhow_muw * flux_w * dI[0] * insideWeight_w;
        dII[2] += fd.tij * up_n;
        dII[3] += rhon_mun * flux_n * dI[1] * insideWeight_n;
        dII[3] += fd.tij * dI[2] * up_n;
    }
}

// flux vector across the same (interior) face seen from the neighbour J (inside), I outside
template <int M>
DMX_HD void face_flux_J(const Phys& ph, const FaceData& fd, const CellState& sI, double* out)
{
    for (int a = 0; a < M; ++a) {
        double F;
        if (ph.gravity) {
            const double rho = (ph.rho[a] + ph.rho[a]) * 0.5;
            F = fd.tij * (fd.o.p[a] - sI.p[a]) + rho * fd.area * fd.bJ;
            F -= rho * fd.tij / fd.ti * (fd.bJ - fd.bI);
        } else {
            F = fd.tij * (fd.o.p[a] - sI.p[a]);
        }
        out[a] = 0.0 + upwind_flux<M>(F, fd.o.up[a], sI.up[a]);
    }
}

template <int DIM, int M>
DMX_HD void local_residual(const Phys& ph, const FaceData* fd, const CellState& sI, const double* SI,
                           const double* sprev, double vol, double dt, bool instationary, double* R)
{
    for (int a = 0; a < M; ++a) {
        double src = 0.0;
        src += 0.0;
        src *= vol * 1.0;
        R[a] = 0.0;
        R[a] -= src;
    }
    for (int f = 0; f < 2 * DIM; ++f) {
        double fl[2];
        face_flux_I<M>(ph, fd[f], sI, fl);
        for (int a = 0; a < M; ++a) R[a] += fl[a];
    }
    if (instationary) {
        for (int a = 0; a < M; ++a) {
            double prev = sprev[a];
            double cur = ph.phi * ph.rho[a] * SI[a];
            prev *= 1.0;
            cur *= 1.0;
            cur -= prev;
            cur *= vol;
            cur /= dt;
            double st = 0.0;
            st += cur;
            R[a] += st;
        }
    }
}

// inverse of a diagonal block (block Jacobi) in FP64, 2x2 closed form; false if singular
template <int M>
DMX_HD bool invert_block(double* D)
{
    if (M == 1) {
        const bool ok = __builtin_fabs(D[0]) > 1e-80;
        D[0] = 1.0 / D[0];
        return ok;
    }
    double detinv = D[0] * D[M * M - 1] - D[1] * D[2 % (M * M)];
    const bool ok = __builtin_fabs(detinv) > 1e-80;
    detinv = 1.0 / detinv;
    const double temp = D[0];
    D[0] = D[M * M - 1] * detinv;
    D[1 % (M * M)] = -D[1 % (M * M)] * detinv;
    D[2 % (M * M)] = -D[2 % (M * M)] * detinv;
    D[M * M - 1] = temp * detinv;
    return ok;
}

template <int M>
DMX_HD void store_block_inverse(const AsmPtrs& P, int64_t I, double* Dii)
{
    if (!P.Dinv32 && !P.Dinv64) return;
    const bool ok = invert_block<M>(Dii);
    if (!ok && P.fail) *P.fail = 1;
    if (P.Dinv32)
        for (int q = 0; q < M * M; ++q) P.Dinv32[I * M * M + q] = float(Dii[q]);
    if (P.Dinv64)
        for (int q = 0; q < M * M; ++q) P.Dinv64[I * M * M + q] = Dii[q];
}

// Residual (JAC = false) or residual + Jacobian of cell I (problem-local). Forward differences
// (AN = false) fill column I (A[I][I], A[J][I]: column ownership); the hand-coded derivatives
// (AN = true) fill row I (A[I][I], A[I][J]: row ownership). Either way every block is written
// by exactly one cell.
template <int DIM, int M, bool JAC, bool AN = false>
DMX_HD void assemble_cell(const GridDesc& g, const Phys& ph, const AsmPtrs& P, int64_t I)
{
    int idx[3];
    idx[0] = int(I % g.n[0]);
    idx[1] = int((I / g.n[0]) % g.n[1]);
    idx[2] = int(I / (int64_t(g.n[0]) * g.n[1]));
    double lo[3], up[3], c[3];
    double vol = 1;
    for (int a = 0; a < DIM; ++a) {
        lo[a] = gcoord(g, a, idx[a]);
        up[a] = gcoord(g, a, idx[a] + 1);
        c[a] = lo[a] + 0.5 * (up[a] - lo[a]);
        vol *= up[a] - lo[a];
    }
    const double KI = P.K[I];
    double uI[2];
    for (int a = 0; a < M; ++a) uI[a] = P.u[I * M + a];
    CellState sI;
    for (int a = 0; a < M; ++a) {
        sI.p[a] = P.vv_p[I * M + a];
        sI.up[a] = P.vv_up[I * M + a];
    }
    double SI[2];
    if (M == 1) SI[0] = 1.0;
    else { SI[1] = uI[1]; SI[0] = 1 - uI[1]; }
    const bool instationary = P.uprev != nullptr;
    double sprev[2] = {0.0, 0.0};
    if (instationary) {
        double Sp[2];
        if (M == 1) Sp[0] = 1.0;
        else { const double Sn = P.uprev[I * M + 1]; Sp[1] = Sn; Sp[0] = 1 - Sn; }
        for (int a = 0; a < M; ++a) sprev[a] = ph.phi * ph.rho[a] * Sp[a];
    }

    FaceData fd[2 * DIM];
    for (int f = 0; f < 2 * DIM; ++f) {
        const int ax = f / 2, s = f % 2;
        FaceData& F = fd[f];
        double area = 1;
        for (int a = 0; a < DIM; ++a)
            if (a != ax) area *= up[a] - lo[a];
        F.area = area;
        const double ipx = s ? up[ax] : lo[ax];
        const double dI = ipx - c[ax];
        const double qI = dI / (dI * dI);
        const double dotI = s ? qI : -qI;
        const double ti = KI * dotI;
        // n.g and -n.g: non-zero only on faces normal to the vertical axis
        const bool vert = ph.gravity && ax == DIM - 1;
        const double ng = vert ? (s ? -9.81 : 9.81) : 0.0;
        const double ngJ = vert ? (s ? 9.81 : -9.81) : 0.0;
        F.aI = KI * ng * 1.0;
        F.bI = KI * ngJ * 1.0;
        const bool boundary = s ? (idx[ax] == g.n[ax] - 1) : (idx[ax] == 0);
        if (!boundary) {
            const int64_t Jc = I + (s ? g.stride[ax] : -g.stride[ax]);
            const int jax = idx[ax] + (s ? 1 : -1);
            const double loJ = gcoord(g, ax, jax);
            const double upJ = gcoord(g, ax, jax + 1);
            const double cJ = loJ + 0.5 * (upJ - loJ);
            const double KJ = P.K[Jc];
            const double dJ = ipx - cJ;
            const double qJ = dJ / (dJ * dJ);
            const double dotJ = s ? qJ : -qJ;
            const double tj = -1.0 * (KJ * dotJ);
            F.kind = 0;
            F.ti = ti;
            F.tj = tj;
            if (ti * tj <= 0.0) F.tij = 0;
            else F.tij = area * (ti * tj) / (ti + tj);
            F.aJ = KJ * ng * 1.0;
            F.bJ = KJ * ngJ * 1.0;
            for (int a = 0; a < M; ++a) {
                F.o.p[a] = P.vv_p[Jc * M + a];
                F.o.up[a] = P.vv_up[Jc * M + a];
            }
            if (JAC) {
                int jdx[3] = {idx[0], idx[1], idx[2]};
                jdx[ax] = jax;
                // I is J's lower neighbour when I sits on J's -ax side (s == 1)
                F.slot = int64_t(P.rowptr[P.rowBase + Jc]) + stencil_pos<DIM>(g, jdx, ax, s ? -1 : 1);
            }
        } else {
            const int64_t b = boundary_face(g, idx, f);
            F.ti = ti;
            F.tj = 0.0;
            F.aJ = 0.0;
            F.bJ = 0.0;
            F.slot = -1;
            if (P.bc_type[b] == 1) {
                F.kind = 1;
                F.tij = area * ti;
                for (int a = 0; a < M; ++a) {
                    F.o.p[a] = P.ghost_p[b * M + a];
                    F.o.up[a] = P.ghost_up[b * M + a];
                }
            } else {
                F.kind = 2;
                F.tij = 0.0;
                for (int a = 0; a < M; ++a) F.gN[a] = P.bc_value[b * M + a];
            }
        }
    }

    double R0[2];
    local_residual<DIM, M>(ph, fd, sI, SI, sprev, vol, P.dt, instationary, R0);

    if (JAC && AN) {
        // CCLocalAssembler<analytic>::assembleJacobianAndResidualImpl (cclocalassembler.hh:513-580)
        double DII[M * M], DIJ[2 * DIM][M * M];
        for (int q = 0; q < M * M; ++q) DII[q] = 0.0;
        if (instationary && M == 2) {
            // TwoPIncompressibleLocalResidual::addStorageDerivatives (2p/incompressiblelocalresidual.hh:93-114)
            const double poreVolume = vol * ph.phi;
            DII[1 % (M * M)] -= poreVolume * ph.rho[0] / P.dt;
            DII[M * M - 1] += poreVolume * ph.rho[1 % M] / P.dt;
        }
        double dI[3] = {0.0, 0.0, 0.0};
        if (M == 2)
            for (int k = 0; k < 3; ++k) dI[k] = P.vv_d[I * 3 + k];
        for (int f = 0; f < 2 * DIM; ++f) {
            for (int q = 0; q < M * M; ++q) DIJ[f][q] = 0.0;
            if (fd[f].kind == 2) continue;   // Neumann: no Robin derivatives for these models
            if (M == 1) {
                // OnePIncompressibleLocalResidual (1p/incompressiblelocalresidual.hh:74-124, 176-192)
                const double upc = ph.rho[0] / ph.mu[0];
                const double deriv = fd[f].tij * upc;
                DII[0] += deriv;
                if (fd[f].kind == 0) DIJ[f][0] -= deriv;
            } else {
                double dO[3] = {0.0, 0.0, 0.0};
                if (fd[f].kind == 0) {
                    const int64_t Jc = I + ((f & 1) ? g.stride[f >> 1] : -g.stride[f >> 1]);
                    for (int k = 0; k < 3; ++k) dO[k] = P.vv_d[Jc * 3 + k];
                }
                twop_flux_derivatives(ph, fd[f], sI, dI, dO, DII, DIJ[f]);
            }
        }
        const int64_t row0 = int64_t(P.rowptr[P.rowBase + I]);
        const int64_t slotII = row0 + stencil_pos<DIM>(g, idx, 0, 0);
        for (int q = 0; q < M * M; ++q) P.J[slotII * M * M + q] = DII[q];
        for (int f = 0; f < 2 * DIM; ++f) {
            if (fd[f].kind != 0) continue;
            const int64_t slot = row0 + stencil_pos<DIM>(g, idx, f >> 1, (f & 1) ? 1 : -1);
            for (int q = 0; q < M * M; ++q) P.J[slot * M * M + q] = DIJ[f][q];
            if (P.S32)
                for (int q = 0; q < M * M; ++q) P.S32[(int64_t(f) * M * M + q) * g.N + I] = float(DIJ[f][q]);
        }
        store_block_inverse<M>(P, I, DII);
    }

    if (JAC && !AN) {
        double N0[2 * DIM][2];
        for (int f = 0; f < 2 * DIM; ++f) {
            if (fd[f].kind != 0) continue;
            double fl[2];
            face_flux_J<M>(ph, fd[f], sI, fl);
            for (int a = 0; a < M; ++a) { N0[f][a] = 0.0; N0[f][a] += fl[a]; }
        }
        const int64_t slotII = int64_t(P.rowptr[P.rowBase + I]) + stencil_pos<DIM>(g, idx, 0, 0);
        double Dii[M * M];
        for (int pv = 0; pv < M; ++pv) {
            double pvd[2];
            for (int a = 0; a < M; ++a) pvd[a] = uI[a];
            const double x0 = pvd[pv];
            const double eps = ph.fdMag[pv] > 0.0 ? ph.fdBase * ph.fdMag[pv] : ph.fdBase * (__builtin_fabs(x0) + 1.0);
            double delta = 0.0;
            delta += eps;
            pvd[pv] = x0 + eps;
            CellState sd;
            double Sd[2];
            volvars<M>(ph, pvd, sd, Sd, P.mat ? P.mat[I] : 0);
            double R1[2];
            local_residual<DIM, M>(ph, fd, sd, Sd, sprev, vol, P.dt, instationary, R1);
            for (int a = 0; a < M; ++a) {
                double d = R1[a];
                d -= R0[a];
                d /= delta;
                double v = 0.0;
                v += d;
                P.J[(slotII * M + a) * M + pv] = v;
                Dii[a * M + pv] = v;
            }
            for (int f = 0; f < 2 * DIM; ++f) {
                if (fd[f].kind != 0) continue;
                double fl[2];
                face_flux_J<M>(ph, fd[f], sd, fl);
                for (int a = 0; a < M; ++a) {
                    double N1 = 0.0;
                    N1 += fl[a];
                    double d = N1;
                    d -= N0[f][a];
                    d /= delta;
                    double v = 0.0;
                    v += d;
                    P.J[(fd[f].slot * M + a) * M + pv] = v;
                    if (P.S32) {
                        // row Jc sees cell I on the opposite side of face f
                        const int64_t Jc = I + ((f & 1) ? g.stride[f >> 1] : -g.stride[f >> 1]);
                        P.S32[(int64_t(f ^ 1) * M * M + a * M + pv) * g.N + Jc] = float(v);
                    }
                }
            }
        }
        store_block_inverse<M>(P, I, Dii);
    }
    for (int a = 0; a < M; ++a) P.r[I * M + a] = R0[a];
}

// undeflected cell state (phase pressures, rho*lambda); with vv_d (2p analytic Jacobian) also
// -dkrw/dsw, -dkrn/dsw, -dpc/dsw at the cell's wetting saturation (the dKr*_dSn, dpc_dSn of
// 2p/incompressiblelocalresidual.hh:212-217)
template <int M>
DMX_HD void cell_state(const Phys& ph, const double* u, double* vv_p, double* vv_up, double* vv_d, const int32_t* mat,
                       int64_t I)
{
    double pv[2];
    for (int a = 0; a < M; ++a) pv[a] = u[I * M + a];
    CellState s;
    double S[2];
    const int mI = mat ? mat[I] : 0;
    volvars<M>(ph, pv, s, S, mI);
    for (int a = 0; a < M; ++a) {
        vv_p[I * M + a] = s.p[a];
        vv_up[I * M + a] = s.up[a];
    }
    if (M == 2 && vv_d) {
        vv_d[I * 3 + 0] = -1.0 * vg_dkrw(ph.law[mI], S[0]);
        vv_d[I * 3 + 1] = -1.0 * vg_dkrn(ph.law[mI], S[0]);
        vv_d[I * 3 + 2] = -1.0 * vg_dpc(ph.law[mI], S[0]);
    }
}

// Dirichlet ghost state of boundary face b (tpfa/elementvolumevariables.hh:326-338), with the
// material of the face's inside cell (bmat, null: material 0)
template <int M>
DMX_HD void ghost_state(const Phys& ph, const int32_t* bc_type, const double* bc_value, const int32_t* bmat,
                        double* ghost_p, double* ghost_up, int64_t b)
{
    if (bc_type[b] != 1) {
        for (int a = 0; a < M; ++a) { ghost_p[b * M + a] = 0.0; ghost_up[b * M + a] = 0.0; }
        return;
    }
    CellState s;
    double S[2];
    volvars<M>(ph, bc_value + b * M, s, S, bmat ? bmat[b] : 0);
    for (int a = 0; a < M; ++a) {
        ghost_p[b * M + a] = s.p[a];
        ghost_up[b * M + a] = s.up[a];
    }
}

} // namespace dmx
