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

}  // namespace asd2
