// Rigorous candidate filter shared by the inner-product passes.
//
// For window j of template t (length m) an inner-product pass delivers
//   dt = sum_i q_i * (x0[j+i] - c)   (c: the pass's centring constant)
// computed in reduced or FP32 arithmetic with |dt - dt_exact| <= eta *
// sqrt(Q2) * sqrt(sum_i (x0[j+i] - c)^2) + eabs.  From the prefix sums this
// function derives an interval [lo, hi] that contains the reference's FP32
// distance d_ref (ZnormSearch.hpp:103-165 with the statistics of
// SkipSequential.cpp:157-179):
//
//   d_exact = Q2 + m - 2 G,  G = (dt - (mu - c) S) / sigma        (exact reals)
//   |d_ref - d_exact| <= E_ref  (FP32 rounding of the reference kernel:
//       D1 (2 sqrt(d) + D1) + gamma_{m/16+36} (sqrt(d) + D1)^2 with
//       D1 = u (3.05 sqrt(m) + 1.03 sqrt(m) |mu| / sigma + sqrt(d)),
//       u = 2^-24; the float casts of mean and std, the rounded reciprocal,
//       the subtraction and the fused product contribute the sqrt(m) terms,
//       the 16 FMA lanes plus the horizontal sum and tail the gamma term)
//
// Windows whose statistics cannot be bounded away from the reference's
// constant-window cut (sigma <= 1e-12) or whose prefix differences are
// swamped by rounding get lo = -inf and hi = +inf: they always become
// candidates and are decided exactly.
#pragma once

#include "tmx_internal.hpp"

namespace tmx {

struct DistInterval {
    double lo, hi;  // bounds on the reference's FP32 distance
    double est;     // point estimate from the fast inner product (NaN if undecidable)
};

__device__ inline DistInterval dist_interval(const Waveform &w, uint64_t j, int m, double Q2, double S, float c,
                                             double dt, double eta, double eabs = 0.0)
{
    constexpr double u = 5.9604644775390625e-08;   // 2^-24
    constexpr double u64 = 1.1102230246251565e-16;  // 2^-53
    const double md = static_cast<double>(m);
    const WindowSums ws = window_sums_fast(w, j, j + m);
    const double S1 = ws.S1, S2 = ws.S2, eS1 = ws.eS1, eS2 = ws.eS2;
    const double V = S2 - S1 * S1 / md;  // m sigma^2
    const double eV = eS2 + (2.0 * fabs(S1) * eS1 + eS1 * eS1) / md + 8.0 * u64 * (fabs(S2) + S1 * S1 / md);
    const double vlo = V - eV;
    // sigma must stay clear of the reference's constant-window cut, 1e-12,
    // and the relative uncertainty of V must be small.
    if (!(vlo > md * 4e-24) || !(eV < 0.25 * V)) {
        DistInterval r;
        r.lo = -INFINITY;
        r.hi = INFINITY;
        r.est = NAN;
        return r;
    }
    const double sig_lo = sqrt(vlo / md) * (1.0 - 4.0 * u64);
    const double sig_hi = sqrt((V + eV) / md) * (1.0 + 4.0 * u64);
    const double mu = S1 / md;
    const double emu = eS1 / md + 2.0 * u64 * fabs(mu);
    const double dmc = fabs(mu - static_cast<double>(c)) + emu;
    const double St2 = (V + eV) + md * dmc * dmc;  // >= sum (x0 - c)^2 over the window
    const double Edot = eta * sqrt(Q2) * sqrt(St2) * (1.0 + 1e-6) + eabs;
    const double N = dt - (mu - static_cast<double>(c)) * S;
    const double eN = Edot + emu * fabs(S) + 4.0 * u64 * (fabs(dt) + fabs((mu - c) * S));
    const double nhi = N + eN, nlo = N - eN;
    const double gmax = nhi >= 0.0 ? nhi / sig_lo : nhi / sig_hi;
    const double gmin = nlo >= 0.0 ? nlo / sig_hi : nlo / sig_lo;
    const double base = Q2 + md;
    double dlo = base - 2.0 * gmax;
    double dhi = base - 2.0 * gmin;
    dlo -= 8.0 * u64 * (base + 2.0 * fabs(gmax));
    dhi += 8.0 * u64 * (base + 2.0 * fabs(gmin));
    // Reference FP32 rounding.
    const double sq = sqrt(dhi > 0.0 ? dhi : 0.0);
    const double rm = sqrt(md);
    const double mu_abs = fabs(mu) + emu;
    const double D1 = (u * 1.001) * (3.05 * rm + 1.03 * rm * mu_abs / sig_lo + sq);
    const double neff = floor(md / 16.0) + 36.0;
    const double gam = neff * u / (1.0 - neff * u);
    const double Eref = (D1 * (2.0 * sq + D1) + gam * (sq + D1) * (sq + D1)) * 1.01;
    DistInterval r;
    r.lo = dlo - Eref;
    r.hi = dhi + Eref;
    r.est = base - 2.0 * N / sqrt(V / md);
    return r;
}

// Relative error factor eta of an FP32 sequential dot product of length m
// with operands rounded once (x0 - c in FP32): gamma_m + u (1 + gamma_m).
__host__ __device__ inline double eta_fp32(int m)
{
    constexpr double u = 5.9604644775390625e-08;
    const double g = m * u / (1.0 - m * u);
    return g + u * (1.0 + g);
}

}  // namespace tmx
