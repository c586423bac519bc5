// Oracle (test infrastructure only): CPU restatement of the regularized van Genuchten law.
//
// Follows, operation for operation:
//   dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:109-118   VanGenuchten::pc
//   dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:151-160   VanGenuchten::dpc_dswe
//   dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:215-225   VanGenuchten::krw
//   dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:237-247   VanGenuchten::dkrw_dswe
//   dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:260-269   VanGenuchten::krn
//   dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:281-292   VanGenuchten::dkrn_dswe
//   dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:419-437   regularized pc
//   dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:505-558   regularized krw / krn and derivatives
//   dumux/material/fluidmatrixinteractions/2p/vangenuchten.hh:566-612   spline / threshold init
//   dumux/material/fluidmatrixinteractions/2p/materiallaw.hh:103-223    TwoPMaterialLaw dispatch
//   dumux/material/fluidmatrixinteractions/2p/efftoabsdefaultpolicy.hh:105-152  sw <-> swe
//   dumux/common/spline.hh:336-347 + dumux/common/splinecommon_.hh:419-476,478-525  2-point spline
// The 2x2 moment system of the spline is solved with the closed-form 2x2 inverse that
// dune-common's DenseMatrix::solve uses for two rows (third-party, not in the reference tree).
#pragma once
#include "detmath.h"

namespace orc {

struct Spline2 {
    double x0, x1, y0, y1, m0, m1; // sampling points and moments

    void set(double xa, double xb, double ya, double yb, double slope0, double slope1)
    {
        x0 = xa; x1 = xb; y0 = ya; y1 = yb;
        const double h = x1 - x0;
        // makeNaturalSystem_ + makeFullSystem_ with two samples: M = [[2,1],[1,2]]
        const double d0 = 6.0 / h * ((y1 - y0) / h - slope0);
        const double d1 = 6.0 / h * (slope1 - (y1 - y0) / h);
        const double M00 = 2.0, M01 = 1.0, M10 = 1.0, M11 = 2.0;
        double detinv = M00 * M11 - M01 * M10;
        detinv = 1.0 / detinv;
        m0 = detinv * (M11 * d0 - M01 * d1);
        m1 = detinv * (M00 * d1 - M10 * d0);
    }

    // splinecommon_.hh:478-503 (segment 0 of a 2-point spline)
    double eval(double x) const
    {
        const double h_i1 = x1 - x0;
        const double x_i = x - x0;
        const double x_i1 = x1 - x;
        const double A_i = (y1 - y0) / h_i1 - h_i1 / 6.0 * (m1 - m0);
        const double B_i = y0 - m0 * (h_i1 * h_i1) / 6.0;
        return m0 * x_i1 * x_i1 * x_i1 / (6.0 * h_i1)
             + m1 * x_i * x_i * x_i / (6.0 * h_i1)
             + A_i * x_i
             + B_i;
    }

    // splinecommon_.hh:507-525
    double evalDerivative(double x) const
    {
        const double h_i1 = x1 - x0;
        const double x_i = x - x0;
        const double x_i1 = x1 - x;
        const double A_i = (y1 - y0) / h_i1 - h_i1 / 6.0 * (m1 - m0);
        return -m0 * x_i1 * x_i1 / (2.0 * h_i1)
             + m1 * x_i * x_i / (2.0 * h_i1)
             + A_i;
    }
};

static inline double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

struct VanGenuchtenLaw {
    // base law (VanGenuchten::Params)
    double alpha, n, m, l;
    // eff-to-abs (TwoPEffToAbsDefaultPolicy::Params)
    double swr, snr;
    // regularization (VanGenuchtenRegularization)
    double pcLowSwe, pcHighSwe, krwHighSwe, krnLowSwe;
    double pcLowSwePcValue, pcHighSwePcValue;
    double pcDerivativeLowSw, pcDerivativeHighSweThreshold, pcDerivativeHighSweEnd;
    Spline2 pcSpline, krwSpline, krnSpline;

    // ---- eff <-> abs -----------------------------------------------------------------
    double swToSwe(double sw) const { return (sw - swr) / (1.0 - swr - snr); }
    double sweToSw(double swe) const { return swe * (1.0 - swr - snr) + swr; }
    double dswe_dsw() const { return 1.0 / (1.0 - swr - snr); }
    double dsw_dswe() const { return 1.0 - swr - snr; }

    // ---- base law (un-regularized), argument swe --------------------------------------
    double basePc(double swe) const
    {
        swe = clamp01(swe);
        return dpow(dpow(swe, -1.0 / m) - 1, 1.0 / n) / alpha;
    }
    double baseDpcDswe(double swe) const
    {
        swe = clamp01(swe);
        const double powSwe = dpow(swe, -1 / m);
        return -1.0 / alpha * dpow(powSwe - 1, 1.0 / n - 1) / n * powSwe / swe / m;
    }
    double baseKrw(double swe) const
    {
        swe = clamp01(swe);
        const double r = 1.0 - dpow(1.0 - dpow(swe, 1.0 / m), m);
        return dpow(swe, l) * r * r;
    }
    double baseDkrwDswe(double swe) const
    {
        swe = clamp01(swe);
        const double x = 1.0 - dpow(swe, 1.0 / m);
        const double xToM = dpow(x, m);
        return (1.0 - xToM) * dpow(swe, l - 1) * ((1.0 - xToM) * l + 2 * xToM * (1.0 - x) / x);
    }
    double baseKrn(double swe) const
    {
        swe = clamp01(swe);
        return dpow(1 - swe, l) * dpow(1 - dpow(swe, 1.0 / m), 2 * m);
    }
    double baseDkrnDswe(double swe) const
    {
        swe = clamp01(swe);
        const double sne = 1.0 - swe;
        const double x = 1.0 - dpow(swe, 1.0 / m);
        return -dpow(sne, l - 1.0) * dpow(x, 2 * m - 1.0) * (l * x + 2.0 * sne / swe * (1.0 - x));
    }

    // ---- TwoPMaterialLaw<false> (absolute sw) ------------------------------------------
    double pcNoReg(double sw) const { return basePc(swToSwe(sw)); }
    double dpcDswNoReg(double sw) const { return baseDpcDswe(swToSwe(sw)) * dswe_dsw(); }
    double krwNoReg(double sw) const { return baseKrw(swToSwe(sw)); }
    double dkrwDswNoReg(double sw) const { return baseDkrwDswe(swToSwe(sw)) * dswe_dsw(); }
    double krnNoReg(double sw) const { return baseKrn(swToSwe(sw)); }
    double dkrnDswNoReg(double sw) const { return baseDkrnDswe(swToSwe(sw)) * dswe_dsw(); }

    // vangenuchten.hh:566-612
    void init(double alpha_, double n_, double l_, double swr_, double snr_,
              double pcLow, double pcHigh, double krnLow, double krwHigh)
    {
        alpha = alpha_; n = n_; m = 1.0 - 1.0 / n_; l = l_;
        swr = swr_; snr = snr_;
        pcLowSwe = pcLow; pcHighSwe = pcHigh; krwHighSwe = krwHigh; krnLowSwe = krnLow;
        {
            const double lowSw = sweToSw(pcLowSwe);
            const double highSw = sweToSw(pcHighSwe);
            const double dsw = dsw_dswe();
            pcDerivativeLowSw = dpcDswNoReg(lowSw) * dsw;
            pcDerivativeHighSweThreshold = dpcDswNoReg(highSw) * dsw;
            pcDerivativeHighSweEnd = 2.0 * (0.0 - pcNoReg(highSw)) / (1.0 - pcHighSwe);
            pcLowSwePcValue = pcNoReg(lowSw);
            pcHighSwePcValue = pcNoReg(highSw);
            if (pcHighSwe < 1.0)
                pcSpline.set(pcHighSwe, 1.0, pcHighSwePcValue, 0, pcDerivativeHighSweThreshold, pcDerivativeHighSweEnd);
        }
        {
            const double lowSw = sweToSw(krnLowSwe);
            const double highSw = sweToSw(krwHighSwe);
            const double dsw = dsw_dswe();
            const double krwHighSw = krwNoReg(highSw);
            const double dkrwHighSw = dkrwDswNoReg(highSw) * dsw;
            const double krnLowSw = krnNoReg(lowSw);
            const double dkrnLowSw = dkrnDswNoReg(lowSw) * dsw;
            if (krwHighSwe < 1.0)
                krwSpline.set(krwHighSwe, 1.0, krwHighSw, 1.0, dkrwHighSw, 0.0);
            if (krnLowSwe > 0.0)
                krnSpline.set(0.0, krnLowSwe, 1.0, krnLowSw, 0.0, dkrnLowSw);
        }
    }

    // ---- regularized law (absolute sw), materiallaw.hh:103-114, 183-194, 211-223 -------
    double pc(double sw) const
    {
        const double swe = swToSwe(sw);
        if (swe <= pcLowSwe) return pcLowSwePcValue + pcDerivativeLowSw * (swe - pcLowSwe);
        else if (swe >= 1.0) return pcDerivativeHighSweEnd * (swe - 1.0);
        else if (swe > pcHighSwe) return pcSpline.eval(swe);
        return basePc(swe);
    }
    double dpcDsw(double sw) const
    {
        const double swe = swToSwe(sw);
        if (swe <= pcLowSwe) return pcDerivativeLowSw * dswe_dsw();
        else if (swe >= 1.0) return pcDerivativeHighSweEnd * dswe_dsw();
        else if (swe > pcHighSwe) return pcSpline.evalDerivative(swe) * dswe_dsw();
        return baseDpcDswe(swe) * dswe_dsw();
    }
    // materiallaw.hh:195-206 + vangenuchten.hh:523-533 (regularized, times dswe/dsw)
    double dkrwDsw(double sw) const
    {
        const double swe = swToSwe(sw);
        if (swe <= 0.0) return 0.0 * dswe_dsw();
        else if (swe >= 1.0) return 0.0 * dswe_dsw();
        else if (swe >= krwHighSwe) return krwSpline.evalDerivative(swe) * dswe_dsw();
        return baseDkrwDswe(swe) * dswe_dsw();
    }
    // materiallaw.hh:229-240 + vangenuchten.hh:553-563
    double dkrnDsw(double sw) const
    {
        const double swe = swToSwe(sw);
        if (swe <= 0.0) return 0.0 * dswe_dsw();
        else if (swe >= 1.0) return 0.0 * dswe_dsw();
        else if (swe <= krnLowSwe) return krnSpline.evalDerivative(swe) * dswe_dsw();
        return baseDkrnDswe(swe) * dswe_dsw();
    }
    double krw(double sw) const
    {
        const double swe = swToSwe(sw);
        if (swe <= 0.0) return 0.0;
        else if (swe >= 1.0) return 1.0;
        else if (swe >= krwHighSwe) return krwSpline.eval(swe);
        return baseKrw(swe);
    }
    double krn(double sw) const
    {
        const double swe = swToSwe(sw);
        if (swe <= 0.0) return 1.0;
        else if (swe >= 1.0) return 0.0;
        else if (swe <= krnLowSwe) return krnSpline.eval(swe);
        return baseKrn(swe);
    }
};

} // namespace orc
