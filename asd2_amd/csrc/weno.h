// weno.h -- HJ-WENO5 derivative shared by the level-set and momentum kernels.
#pragma once

#include "internal.h"

namespace asd2 {

// Jiang-Peng HJ-WENO5 one-sided derivatives (reference UnitHJWENO5,
// src/levelset2d.cpp:159-262) from the six undivided forward differences
// D0..D5 = F[i-2]-F[i-3], ..., F[i+3]-F[i+2].  The smoothness indicators are scaled
// by d^2 (so the weight epsilon becomes eps*d^2) and the three weights of a side share
// one division.
__device__ __forceinline__ double weno_side(double is0, double is1, double is2, double q0,
                                            double q1, double q2, double epsd2) {
    const double s0 = (epsd2 + is0) * (epsd2 + is0);
    const double s1 = (epsd2 + is1) * (epsd2 + is1);
    const double s2 = (epsd2 + is2) * (epsd2 + is2);
    const double t0 = 0.1 * (s1 * s2);
    const double t1 = 0.6 * (s0 * s2);
    const double t2 = 0.3 * (s0 * s1);
    return (t0 * q0 + t1 * q1 + t2 * q2) / (t0 + t1 + t2);
}

__device__ __forceinline__ void weno5(double D0, double D1, double D2, double D3, double D4, double D5,
                                      double inv_d, double epsd2, double& fp, double& fm) {
    constexpr double c13 = 13.0 / 12.0;
    const double A1 = D0 - 2.0 * D1 + D2;
    const double A2 = D1 - 2.0 * D2 + D3;
    const double A3 = D2 - 2.0 * D3 + D4;
    const double A4 = D3 - 2.0 * D4 + D5;
    const double b0 = D0 - 4.0 * D1 + 3.0 * D2;
    const double b1 = D1 - D3;
    const double b2 = 3.0 * D2 - 4.0 * D3 + D4;
    const double e0 = D5 - 4.0 * D4 + 3.0 * D3;
    const double e1 = D4 - D2;
    const double e2 = 3.0 * D3 - 4.0 * D2 + D1;
    const double qm0 = (1.0 / 3.0) * D0 - (7.0 / 6.0) * D1 + (11.0 / 6.0) * D2;
    const double qa = (-1.0 / 6.0) * D1 + (5.0 / 6.0) * D2 + (1.0 / 3.0) * D3;
    const double qb = (1.0 / 3.0) * D2 + (5.0 / 6.0) * D3 - (1.0 / 6.0) * D4;
    const double qp0 = (1.0 / 3.0) * D5 - (7.0 / 6.0) * D4 + (11.0 / 6.0) * D3;
    fm = inv_d * weno_side(c13 * A1 * A1 + 0.25 * b0 * b0, c13 * A2 * A2 + 0.25 * b1 * b1,
                           c13 * A3 * A3 + 0.25 * b2 * b2, qm0, qa, qb, epsd2);
    fp = inv_d * weno_side(c13 * A4 * A4 + 0.25 * e0 * e0, c13 * A3 * A3 + 0.25 * e1 * e1,
                           c13 * A2 * A2 + 0.25 * e2 * e2, qp0, qb, qa, epsd2);
}

// The same derivatives in the reference's own operation order (src/levelset2d.cpp:159-262: every
// smoothness term divided by d, the three weights alpha_k = c_k / (eps + IS_k)^2 and
// w_k = alpha_k / (alpha_0 + alpha_1 + alpha_2) each by its own division, each stencil term divided
// by d), with contraction off: the result is the reference's to the last bit.  POW2: d is a power of
// two, so x / d equals x * (1 / d) exactly and the divisions by d become multiplications.
template <bool POW2>
__device__ __forceinline__ double wdiv(double x, double d, double inv_d) {
    return POW2 ? x * inv_d : x / d;
}

// Correctly rounded quotients a / b by a Newton-Raphson reciprocal and Markstein's correction
// (P. Markstein, "Computation of elementary functions on the IBM RISC System/6000 processor", IBM J.
// Res. Dev. 34, 1990): y0 = the hardware reciprocal estimate, two Newton steps y' = y + y (1 - b y),
// then q = a y, q' = q + (a - b q) y.  This is the division operator's own sequence on this chip
// without its range-scaling steps, which are identities while a, b, 1 / b and a / b are normal and
// far from the range limits.  The smoothness weights below stay there: the alphas' divisors
// (1e-6 + IS)^2 are at least 1e-12 and the alphas at most 6e11; the weights' divisor
// alpha_0 + alpha_1 + alpha_2 is at least each alpha, so the weights are at most 1; all of them are
// normal numbers far from the limits while the smoothness indicators IS stay below about 1e100,
// which bounded level-set fluxes keep by a wide margin.  One reciprocal serves every quotient with the
// same divisor.
struct NrRecip {
    double b, y;
};
__device__ __forceinline__ NrRecip nr_recip(double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    return NrRecip{b, y};
}
__device__ __forceinline__ double nr_div(double a, const NrRecip& r) {
    const double q = a * r.y;
    const double rem = fma(-r.b, q, a);
    return fma(rem, r.y, q);
}

template <bool POW2>
__device__ __forceinline__ double weno_ref_side(double Da, double Db, double Dc, double Dd, double De, double d,
                                                double inv_d) {
    // Da..De = D[i-3]..D[i+1] (the minus side); the plus side passes D[i+2]..D[i-2]
#pragma clang fp contract(off)
    const double t0 = wdiv<POW2>(Da - 2.0 * Db + Dc, d, inv_d), u0 = wdiv<POW2>(Da - 4.0 * Db + 3.0 * Dc, d, inv_d);
    const double t1 = wdiv<POW2>(Db - 2.0 * Dc + Dd, d, inv_d), u1 = wdiv<POW2>(Db - Dd, d, inv_d);
    const double t2 = wdiv<POW2>(Dc - 2.0 * Dd + De, d, inv_d), u2 = wdiv<POW2>(3.0 * Dc - 4.0 * Dd + De, d, inv_d);
    const double IS0 = 13.0 / 12.0 * (t0 * t0) + 1.0 / 4.0 * (u0 * u0);
    const double IS1 = 13.0 / 12.0 * (t1 * t1) + 1.0 / 4.0 * (u1 * u1);
    const double IS2 = 13.0 / 12.0 * (t2 * t2) + 1.0 / 4.0 * (u2 * u2);
    const double s0 = 1e-6 + IS0, s1 = 1e-6 + IS1, s2 = 1e-6 + IS2;
    const double a0 = nr_div(0.1, nr_recip(s0 * s0)), a1 = nr_div(0.6, nr_recip(s1 * s1));
    const double a2 = nr_div(0.3, nr_recip(s2 * s2));
    const NrRecip rs = nr_recip(a0 + a1 + a2);
    const double w0 = nr_div(a0, rs), w1 = nr_div(a1, rs), w2 = nr_div(a2, rs);
    const double q0 = 1.0 / 3.0 * Da - 7.0 / 6.0 * Db + 11.0 / 6.0 * Dc;
    const double q1 = -1.0 / 6.0 * Db + 5.0 / 6.0 * Dc + 1.0 / 3.0 * Dd;
    const double q2 = 1.0 / 3.0 * Dc + 5.0 / 6.0 * Dd - 1.0 / 6.0 * De;
    return wdiv<POW2>(w0 * q0, d, inv_d) + wdiv<POW2>(w1 * q1, d, inv_d) + wdiv<POW2>(w2 * q2, d, inv_d);
}

// D0..D5 = D[i-3]..D[i+2] with D[k] = F[k+1] - F[k]
template <bool POW2>
__device__ __forceinline__ void weno5_ref(double D0, double D1, double D2, double D3, double D4, double D5, double d,
                                          double inv_d, double& fp, double& fm) {
    fm = weno_ref_side<POW2>(D0, D1, D2, D3, D4, d, inv_d);
    fp = weno_ref_side<POW2>(D5, D4, D3, D2, D1, d, inv_d);
}

// d is a power of two (then 1 / d is exact)
inline bool is_pow2(double d) {
    int e = 0;
    return d > 0.0 && std::frexp(d, &e) == 0.5;
}

}  // namespace asd2
