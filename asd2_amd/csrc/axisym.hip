// axisym.hip -- the axisymmetric (r, z) pieces of MACSolver2DAxisym that are not shared with the
// planar solver: the curvature (src/mac2dAxisym.cpp:222-262), the two divergences (:1974-2008)
// and the tridiagonal solve
// of its semi-implicit line sweeps (SolveTDMA, :2255-2304), batched.  The pressure system comes
// from the ghost-fluid assembly kernel in its axisymmetric mode (gfm.hip), the projection from
// the planar UpdateVel kernel (the reference's axisymmetric UpdateVel, :2010-2092, computes the
// same face updates), the level set from the planar level-set kernels.  x = r (base radius =
// the grid's base x), y = z.
#include "internal.h"

namespace asd2 {
namespace {

// FOR_POISSON: (r_{i+1/2} u_{i+1} - r_{i-1/2} u_i) / dr + r_i (v_{j+1} - v_j) / dz
// else:        (u_i + u_{i+1}) / (2 r_i) + (u_{i+1} - u_i) / dr + (v_{j+1} - v_j) / dz
// zero outside the interior
template <bool FOR_POISSON>
__global__ __launch_bounds__(256) void k_div_axisym(Geom G, double base_r, const double* __restrict__ u,
                                                    const double* __restrict__ v, double* __restrict__ div) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + 2 * G.g || j >= G.ny + 2 * G.g) return;
    double d = 0.0;
    if (i >= G.g && i < G.nx + G.g && j >= G.g && j < G.ny + G.g) {
        const double rPM = base_r + (i + 0.5 - G.g) * G.dx;
        const double uW = u[G.idx(i, j)], uE = u[G.idx(i + 1, j)];
        const double vS = v[G.idx(i, j)], vN = v[G.idx(i, j + 1)];
        if (FOR_POISSON) {
            const double rPEHalf = base_r + (i + 1.0 - G.g) * G.dx, rPWHalf = base_r + (i - G.g) * G.dx;
            d = (rPEHalf * uE - rPWHalf * uW) / G.dx + rPM * (vN - vS) / G.dy;
        } else {
            d = 0.5 * (uW + uE) / rPM + (uE - uW) / G.dx + (vN - vS) / G.dy;
        }
    }
    div[G.idx(i, j)] = d;
}

// d phi / dz as the reference stores it: central difference on the interior, 0 elsewhere
__device__ __forceinline__ double dz_at(const Geom& G, const double* __restrict__ ls, int i, int j) {
    if (i < G.g || i >= G.nx + G.g) return 0.0;
    return (ls[G.idx(i, j + 1)] - ls[G.idx(i, j - 1)]) / (2.0 * G.dy);
}

// MACSolver2DAxisym::UpdateKappa (src/mac2dAxisym.cpp:222-262): the planar curvature plus the hoop
// term phi_r |grad phi|^2 / r, the r derivative one-sided on the axis row (i = g), then the clip
// to 1 / min(dr, dz); interior cells only
__global__ __launch_bounds__(256) void k_kappa_axisym(Geom G, double base_r, const double* __restrict__ ls,
                                                      double* __restrict__ kap) {
#pragma clang fp contract(off)
    const int j = G.g + blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = G.g + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + G.g || j >= G.ny + G.g) return;
    const int64_t c = G.idx(i, j);
    const double dr = G.dx, dz = G.dy;
    const double lW = ls[G.idx(i - 1, j)], lE = ls[G.idx(i + 1, j)], lM = ls[c];
    const double dR = (i == G.g) ? (lE - lM) / dr : (lE - lW) / (2.0 * dr);
    const double dZ = dz_at(G, ls, i, j);
    const double S = sqrt(dR * dR + dZ * dZ + 1.0e-100);
    const double rPM = base_r + (i + 0.5 - G.g) * dr;
    const double k = ((lW - 2.0 * lM + lE) / (dr * dr) * dZ * dZ
                      - 2.0 * (dz_at(G, ls, i + 1, j) - dz_at(G, ls, i - 1, j)) / (2.0 * dr) * dR * dZ
                      + (ls[G.idx(i, j - 1)] - 2.0 * lM + ls[G.idx(i, j + 1)]) / (dz * dz) * dR * dR
                      + dR * pow(S, 2.0) / rPM) / pow(S, 3.0);
    kap[c] = isign(k) * fabs(smin(fabs(k), 1.0 / smin(dr, dz)));
}

// one thread per system; element k of system s at k * count + s (coalesced across systems).
// The reference's recurrence: c_0 /= b_0, d_0 /= b_0; c_k /= b_k - a_k c_{k-1},
// d_k = (d_k - a_k d_{k-1}) / (b_k - a_k c_{k-1}); back substitution d_k -= c_k d_{k+1}.
__global__ __launch_bounds__(256) void k_tdma(int n, int count, const double* __restrict__ a,
                                              const double* __restrict__ b, double* __restrict__ c,
                                              double* __restrict__ d) {
#pragma clang fp contract(off)
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= count) return;
    auto at = [&](int k) { return static_cast<int64_t>(k) * count + s; };
    const int m = n - 1;
    c[at(0)] /= b[at(0)];
    d[at(0)] /= b[at(0)];
    for (int k = 1; k < m; k++) {
        c[at(k)] /= b[at(k)] - a[at(k)] * c[at(k - 1)];
        d[at(k)] = (d[at(k)] - a[at(k)] * d[at(k - 1)]) / (b[at(k)] - a[at(k)] * c[at(k - 1)]);
    }
    d[at(m)] = (d[at(m)] - a[at(m)] * d[at(m - 1)]) / (b[at(m)] - a[at(m)] * c[at(m - 1)]);
    for (int k = m; k-- > 0;) d[at(k)] -= c[at(k)] * d[at(k + 1)];
}

}  // namespace
}  // namespace asd2

using namespace asd2;

int asd2_mac_divergence_axisym(asd2_ctx* ctx, const double* d_u, const double* d_v, double* d_div, int for_poisson) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_u || !d_v || !d_div) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    const dim3 grid(cdiv(G.ny + 2 * G.g, 64), cdiv(G.nx + 2 * G.g, 4));
    if (for_poisson)
        hipLaunchKernelGGL(k_div_axisym<true>, grid, dim3(256), 0, ctx->stream, G, ctx->base_x, d_u, d_v, d_div);
    else
        hipLaunchKernelGGL(k_div_axisym<false>, grid, dim3(256), 0, ctx->stream, G, ctx->base_x, d_u, d_v, d_div);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_mac_kappa_axisym(asd2_ctx* ctx, const double* d_ls, double* d_kappa) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_kappa) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    hipLaunchKernelGGL(k_kappa_axisym, dim3(cdiv(G.ny, 64), cdiv(G.nx, 4)), dim3(256), 0, ctx->stream, G,
                       ctx->base_x, d_ls, d_kappa);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_tdma(asd2_ctx* ctx, int n, int count, const double* d_a, const double* d_b, double* d_c, double* d_d) {
    ASD2_TRY(check_ctx(ctx));
    if (n < 2 || count < 1 || !d_a || !d_b || !d_c || !d_d) {
        set_error("asd2_tdma: needs n >= 2 unknowns, count >= 1 systems and four arrays");
        return ASD2_ERR_ARG;
    }
    hipLaunchKernelGGL(k_tdma, dim3(cdiv(count, 256)), dim3(256), 0, ctx->stream, n, count, d_a, d_b, d_c, d_d);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}
