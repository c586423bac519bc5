// Constitutive relations of the OneP / TwoP (pw-Sn) immiscible CCTpfa path, host + device.
//
// Reference arithmetic (same operation order, so the device residual equals the host one
// bit for bit when built with -ffp-contract=off):
//   van Genuchten base law     dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:109-292
//   regularization + splines   dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:419-612
//   TwoPMaterialLaw dispatch   dumux/material/fluidmatrixinteractions/2p/materiallaw.hh:103-223
//   eff <-> abs saturation     dumux/material/fluidmatrixinteractions/2p/efftoabsdefaultpolicy.hh:105-152
//   2-point spline             dumux/common/spline.hh:336-347, dumux/common/splinecommon_.hh:419-525
//   volume variables           dumux/porousmediumflow/2p/volumevariables.hh:74-191,
//                              dumux/porousmediumflow/1p/volumevariables.hh:64-120,178-179
//   storage                    dumux/porousmediumflow/immiscible/localresidual.hh:64-84
#pragma once
#include "detmath.hh"

namespace dmx {

struct Spline2 {
    double x0, x1, y0, y1, m0, m1;
};

// 2-point cubic spline with prescribed end slopes; the 2x2 moment system [[2,1],[1,2]] is
// solved with the closed-form 2x2 inverse (dune-common DenseMatrix::solve for two rows).
inline void spline_set(Spline2& s, double xa, double xb, double ya, double yb, double slope0, double slope1)
{
    s.x0 = xa; s.x1 = xb; s.y0 = ya; s.y1 = yb;
    const double h = s.x1 - s.x0;
    const double d0 = 6.0 / h * ((s.y1 - s.y0) / h - slope0);
    const double d1 = 6.0 / h * (slope1 - (s.y1 - s.y0) / h);
    const double M00 = 2.0, M01 = 1.0, M10 = 1.0, M11 = 2.0;
    double detinv = M00 * M11 - M01 * M10;
    detinv = 1.0 / detinv;
    s.m0 = detinv * (M11 * d0 - M01 * d1);
    s.m1 = detinv * (M00 * d1 - M10 * d0);
}

DMX_HD double spline_eval(const Spline2& s, double x)
{
    const double h_i1 = s.x1 - s.x0;
    const double x_i = x - s.x0;
    const double x_i1 = s.x1 - x;
    const double A_i = (s.y1 - s.y0) / h_i1 - h_i1 / 6.0 * (s.m1 - s.m0);
    const double B_i = s.y0 - s.m0 * (h_i1 * h_i1) / 6.0;
    return s.m0 * x_i1 * x_i1 * x_i1 / (6.0 * h_i1)
         + s.m1 * x_i * x_i * x_i / (6.0 * h_i1)
         + A_i * x_i
         + B_i;
}

DMX_HD double spline_deriv(const Spline2& s, double x)
{
    const double h_i1 = s.x1 - s.x0;
    const double x_i = x - s.x0;
    const double x_i1 = s.x1 - x;
    const double A_i = (s.y1 - s.y0) / h_i1 - h_i1 / 6.0 * (s.m1 - s.m0);
    return -s.m0 * x_i1 * x_i1 / (2.0 * h_i1) + s.m1 * x_i * x_i / (2.0 * h_i1) + A_i;
}

// VanGenuchtenDefault = TwoPMaterialLaw<VanGenuchten, VanGenuchtenRegularization, TwoPEffToAbsDefaultPolicy>
struct VGLaw {
    double alpha, n, m, l;
    double swr, snr;
    double pcLowSwe, pcHighSwe, krwHighSwe, krnLowSwe;
    double pcLowSwePcValue, pcHighSwePcValue;
    double pcDerivativeLowSw, pcDerivativeHighSweThreshold, pcDerivativeHighSweEnd;
    Spline2 pcSpline, krwSpline, krnSpline;
};

DMX_HD double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }
DMX_HD double vg_swToSwe(const VGLaw& L, double sw) { return (sw - L.swr) / (1.0 - L.swr - L.snr); }
DMX_HD double vg_sweToSw(const VGLaw& L, double swe) { return swe * (1.0 - L.swr - L.snr) + L.swr; }

DMX_HD double vg_base_pc(const VGLaw& L, double swe)
{
    swe = clamp01(swe);
    return dmx::pow(dmx::pow(swe, -1.0 / L.m) - 1, 1.0 / L.n) / L.alpha;
}
DMX_HD double vg_base_dpc(const VGLaw& L, double swe)
{
    swe = clamp01(swe);
    const double powSwe = dmx::pow(swe, -1 / L.m);
    return -1.0 / L.alpha * dmx::pow(powSwe - 1, 1.0 / L.n - 1) / L.n * powSwe / swe / L.m;
}
DMX_HD double vg_base_krw(const VGLaw& L, double swe)
{
    swe = clamp01(swe);
    const double r = 1.0 - dmx::pow(1.0 - dmx::pow(swe, 1.0 / L.m), L.m);
    return dmx::pow(swe, L.l) * r * r;
}
DMX_HD double vg_base_dkrw(const VGLaw& L, double swe)
{
    swe = clamp01(swe);
    const double x = 1.0 - dmx::pow(swe, 1.0 / L.m);
    const double xToM = dmx::pow(x, L.m);
    return (1.0 - xToM) * dmx::pow(swe, L.l - 1) * ((1.0 - xToM) * L.l + 2 * xToM * (1.0 - x) / x);
}
DMX_HD double vg_base_krn(const VGLaw& L, double swe)
{
    swe = clamp01(swe);
    return dmx::pow(1 - swe, L.l) * dmx::pow(1 - dmx::pow(swe, 1.0 / L.m), 2 * L.m);
}
DMX_HD double vg_base_dkrn(const VGLaw& L, double swe)
{
    swe = clamp01(swe);
    const double sne = 1.0 - swe;
    const double x = 1.0 - dmx::pow(swe, 1.0 / L.m);
    return -dmx::pow(sne, L.l - 1.0) * dmx::pow(x, 2 * L.m - 1.0) * (L.l * x + 2.0 * sne / swe * (1.0 - x));
}

// vangenuchten.hh:566-612, evaluated once per material on the host
inline void vg_init(VGLaw& L, double alpha, double n, double l, double swr, double snr,
                    double pcLow, double pcHigh, double krnLow, double krwHigh)
{
    L.alpha = alpha; L.n = n; L.m = 1.0 - 1.0 / n; L.l = l;
    L.swr = swr; L.snr = snr;
    L.pcLowSwe = pcLow; L.pcHighSwe = pcHigh; L.krwHighSwe = krwHigh; L.krnLowSwe = krnLow;
    const double dsw_dswe = 1.0 - L.swr - L.snr;
    const double dswe_dsw = 1.0 / (1.0 - L.swr - L.snr);
    {
        const double lowSw = vg_sweToSw(L, L.pcLowSwe);
        const double highSw = vg_sweToSw(L, L.pcHighSwe);
        L.pcDerivativeLowSw = (vg_base_dpc(L, vg_swToSwe(L, lowSw)) * dswe_dsw) * dsw_dswe;
        L.pcDerivativeHighSweThreshold = (vg_base_dpc(L, vg_swToSwe(L, highSw)) * dswe_dsw) * dsw_dswe;
        L.pcDerivativeHighSweEnd = 2.0 * (0.0 - vg_base_pc(L, vg_swToSwe(L, highSw))) / (1.0 - L.pcHighSwe);
        L.pcLowSwePcValue = vg_base_pc(L, vg_swToSwe(L, lowSw));
        L.pcHighSwePcValue = vg_base_pc(L, vg_swToSwe(L, highSw));
        if (L.pcHighSwe < 1.0)
            spline_set(L.pcSpline, L.pcHighSwe, 1.0, L.pcHighSwePcValue, 0, L.pcDerivativeHighSweThreshold,
                       L.pcDerivativeHighSweEnd);
    }
    {
        const double lowSw = vg_sweToSw(L, L.krnLowSwe);
        const double highSw = vg_sweToSw(L, L.krwHighSwe);
        const double krwHighSw = vg_base_krw(L, vg_swToSwe(L, highSw));
        const double dkrwHighSw = (vg_base_dkrw(L, vg_swToSwe(L, highSw)) * dswe_dsw) * dsw_dswe;
        const double krnLowSw = vg_base_krn(L, vg_swToSwe(L, lowSw));
        const double dkrnLowSw = (vg_base_dkrn(L, vg_swToSwe(L, lowSw)) * dswe_dsw) * dsw_dswe;
        if (L.krwHighSwe < 1.0) spline_set(L.krwSpline, L.krwHighSwe, 1.0, krwHighSw, 1.0, dkrwHighSw, 0.0);
        if (L.krnLowSwe > 0.0) spline_set(L.krnSpline, 0.0, L.krnLowSwe, 1.0, krnLowSw, 0.0, dkrnLowSw);
    }
}

DMX_HD double vg_pc(const VGLaw& L, double sw)
{
    const double swe = vg_swToSwe(L, sw);
    if (swe <= L.pcLowSwe) return L.pcLowSwePcValue + L.pcDerivativeLowSw * (swe - L.pcLowSwe);
    else if (swe >= 1.0) return L.pcDerivativeHighSweEnd * (swe - 1.0);
    else if (swe > L.pcHighSwe) return spline_eval(L.pcSpline, swe);
    return vg_base_pc(L, swe);
}

DMX_HD double vg_krw(const VGLaw& L, double sw)
{
    const double swe = vg_swToSwe(L, sw);
    if (swe <= 0.0) return 0.0;
    else if (swe >= 1.0) return 1.0;
    else if (swe >= L.krwHighSwe) return spline_eval(L.krwSpline, swe);
    return vg_base_krw(L, swe);
}

DMX_HD double vg_krn(const VGLaw& L, double sw)
{
    const double swe = vg_swToSwe(L, sw);
    if (swe <= 0.0) return 1.0;
    else if (swe >= 1.0) return 0.0;
    else if (swe <= L.krnLowSwe) return spline_eval(L.krnSpline, swe);
    return vg_base_krn(L, swe);
}

// Derivatives of TwoPMaterialLaw w.r.t. the absolute wetting saturation (materiallaw.hh:117-129,
// 195-206, 229-240): the regularization's branch derivative (vangenuchten.hh:441-455, 523-533,
// 553-563) or the base law's, times dswe/dsw (efftoabsdefaultpolicy.hh).
DMX_HD double vg_dpc(const VGLaw& L, double sw)
{
    const double swe = vg_swToSwe(L, sw);
    const double dswe_dsw = 1.0 / (1.0 - L.swr - L.snr);
    if (swe <= L.pcLowSwe) return L.pcDerivativeLowSw * dswe_dsw;
    else if (swe >= 1.0) return L.pcDerivativeHighSweEnd * dswe_dsw;
    else if (swe > L.pcHighSwe) return spline_deriv(L.pcSpline, swe) * dswe_dsw;
    return vg_base_dpc(L, swe) * dswe_dsw;
}

DMX_HD double vg_dkrw(const VGLaw& L, double sw)
{
    const double swe = vg_swToSwe(L, sw);
    const double dswe_dsw = 1.0 / (1.0 - L.swr - L.snr);
    if (swe <= 0.0) return 0.0 * dswe_dsw;
    else if (swe >= 1.0) return 0.0 * dswe_dsw;
    else if (swe >= L.krwHighSwe) return spline_deriv(L.krwSpline, swe) * dswe_dsw;
    return vg_base_dkrw(L, swe) * dswe_dsw;
}

DMX_HD double vg_dkrn(const VGLaw& L, double sw)
{
    const double swe = vg_swToSwe(L, sw);
    const double dswe_dsw = 1.0 / (1.0 - L.swr - L.snr);
    if (swe <= 0.0) return 0.0 * dswe_dsw;
    else if (swe >= 1.0) return 0.0 * dswe_dsw;
    else if (swe <= L.krnLowSwe) return spline_deriv(L.krnSpline, swe) * dswe_dsw;
    return vg_base_dkrn(L, swe) * dswe_dsw;
}

// Fluid/solid constants of one problem (incompressible phases, constant porosity).
struct Phys {
    int m;            // 1: OneP, 2: TwoP
    int gravity;
    double rho[2], mu[2];
    double phi;
    double fdBase;       // NumericEpsilon base epsilon
    double fdMag[2];     // NumericEpsilon primary-variable magnitudes (<= 0: |x|+1)
    VGLaw law[2];        // material 0 and material 1 (element-wise materials)
};

// Per-cell state the fluxes need: phase pressures and upwind terms rho*lambda.
struct CellState {
    double p[2];
    double up[2];
};

// mat: material of the element whose spatial params apply (0 or 1)
template <int M>
DMX_HD void volvars(const Phys& ph, const double* pv, CellState& s, double* S, int mat)
{
    if (M == 1) {
        s.p[0] = pv[0];
        s.up[0] = ph.rho[0] * (1.0 / ph.mu[0]);
        S[0] = 1.0;
    } else {
        const double Sn = pv[1];
        S[1] = Sn;
        S[0] = 1 - Sn;
        const VGLaw& L = ph.law[mat];
        const double pc = vg_pc(L, S[0]);
        s.p[0] = pv[0];
        s.p[1] = pv[0] + pc;
        s.up[0] = ph.rho[0] * (vg_krw(L, S[0]) / ph.mu[0]);
        s.up[1] = ph.rho[1] * (vg_krn(L, S[0]) / ph.mu[1]);
    }
}

} // namespace dmx
