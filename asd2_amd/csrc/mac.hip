// mac.hip -- staggered-grid pieces of the two-phase step around the pressure solve
// (reference MACSolver2D, src/mac2d.cpp): curvature, Heaviside functions, divergence,
// GFM-consistent projection, CFL time step, and the momentum predictor.
#include <algorithm>
#include <cmath>
#include <limits>

#include "internal.h"
#include "weno.h"

namespace asd2 {

inline dim3 rows_grid(const Geom& G) { return dim3(cdiv(G.ny + 2 * G.g, 64), cdiv(G.nx + 2 * G.g, 4)); }

// ------------------------------------------------------------------ curvature
// src/mac2d.cpp:221-252: central differences; phi_xy from differences of the stored
// phi_x at j +- 1, which is 0 outside the interior rows (reference boundary artefact).
__device__ __forceinline__ double dlsdx_at(const Geom& G, const double* ls, int i, int j) {
#pragma clang fp contract(off)
    if (j < G.g || j >= G.ny + G.g) return 0.0;
    return (ls[G.idx(i + 1, j)] - ls[G.idx(i - 1, j)]) / (2.0 * G.dx);
}

__global__ __launch_bounds__(256) void k_kappa(Geom G, const double* __restrict__ ls, double* __restrict__ kap) {
#pragma clang fp contract(off)
    const int j = G.g + blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = G.g + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + G.g || j >= G.ny + G.g) return;
    const int64_t c = G.idx(i, j);
    const double dx = G.dx, dy = G.dy;
    const double dX = (ls[G.idx(i + 1, j)] - ls[G.idx(i - 1, j)]) / (2.0 * dx);
    const double dY = (ls[G.idx(i, j + 1)] - ls[G.idx(i, j - 1)]) / (2.0 * dy);
    const double S = sqrt(dX * dX + dY * dY + 1.0e-100);
    const double k = (dX * dX * (ls[G.idx(i, j - 1)] - 2.0 * ls[c] + ls[G.idx(i, j + 1)]) / (dy * dy)
                    - 2.0 * dX * dY * (dlsdx_at(G, ls, i, j + 1) - dlsdx_at(G, ls, i, j - 1)) / (2.0 * dy)
                    + dY * dY * (ls[G.idx(i - 1, j)] - 2.0 * ls[c] + ls[G.idx(i + 1, j)]) / (dx * dx))
                   / pow(S, 3.0);
    kap[c] = isign(k) * fabs(smin(fabs(k), 1.0 / smin(dx, dy)));
}

// src/mac2d.cpp:87-122 over the padded array
__global__ __launch_bounds__(256) void k_heaviside(int64_t n, double eps, int smooth, const double* __restrict__ ls,
                                                   double* __restrict__ H) {
#pragma clang fp contract(off)
    for (int64_t k = blockIdx.x * 256ll + threadIdx.x; k < n; k += static_cast<int64_t>(gridDim.x) * 256) {
        const double x = ls[k];
        double h;
        if (!smooth) h = (x >= 0.0) ? 1.0 : 0.0;
        else if (x > eps) h = 1.0;
        else if (x < -eps) h = 0.0;
        else h = 0.5 * (1.0 + x / eps + 1.0 / M_PI * sin(M_PI * x / eps));
        H[k] = h;
    }
}

// src/mac2d.cpp:1729-1738 (zero outside the interior)
__global__ __launch_bounds__(256) void k_divergence(Geom G, const double* __restrict__ u, const double* __restrict__ v,
                                                    double* __restrict__ div) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + 2 * G.g || j >= G.ny + 2 * G.g) return;
    double d = 0.0;
    if (i >= G.g && i < G.nx + G.g && j >= G.g && j < G.ny + G.g)
        d = (u[G.idx(i + 1, j)] - u[G.idx(i, j)]) / G.dx + (v[G.idx(i, j + 1)] - v[G.idx(i, j)]) / G.dy;
    div[G.idx(i, j)] = d;
}

// p = ps / dt on the interior, the ghosts of p kept (src/mac2d.cpp:1722-1724)
__global__ __launch_bounds__(256) void k_pressure(Geom G, const double* __restrict__ ps, double* __restrict__ p,
                                                  double dt) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i < G.g || i >= G.nx + G.g || j < G.g || j >= G.ny + G.g) return;
    p[G.idx(i, j)] = ps[G.idx(i, j)] / dt;
}

// ------------------------------------------------------------------ projection
// face between a (west/south cell) and b (the cell itself), src/mac2d.cpp:1775-1788
__device__ __forceinline__ void face_theta_w(double a, double b, double& thH, double& th) {
#pragma clang fp contract(off)
    if (a >= 0.0 && b >= 0.0) { thH = 1.0; th = 0.0; }
    else if (a < 0.0 && b < 0.0) { thH = 0.0; th = 0.0; }
    else if (a >= 0.0 && b < 0.0) { thH = fabs(a) / (fabs(a) + fabs(b)); th = fabs(a) / (fabs(a) + fabs(b)); }
    else { thH = fabs(b) / (fabs(a) + fabs(b)); th = fabs(a) / (fabs(a) + fabs(b)); }
}

struct ProjArgs {
    double rho_h, rho_l, sigma, dt;
};

// u on faces i in [g+1, nx+g), v on faces j in [g+1, ny+g); src/mac2d.cpp:1768-1843
__global__ __launch_bounds__(256) void k_update_vel(Geom G, ProjArgs A, double* __restrict__ u, double* __restrict__ v,
                                                    const double* __restrict__ us, const double* __restrict__ vs,
                                                    const double* __restrict__ ps, const double* __restrict__ ls,
                                                    const double* __restrict__ kap, const double* __restrict__ H) {
#pragma clang fp contract(off)
    const int j = G.g + blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = G.g + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + G.g || j >= G.ny + G.g) return;
    const int64_t c = G.idx(i, j);
    if (i >= G.g + 1) {
        const int64_t w = G.idx(i - 1, j);
        double thH, th;
        face_theta_w(ls[w], ls[c], thH, th);
        const double iRho = 1.0 / (A.rho_h * thH + A.rho_l * (1.0 - thH));
        const double kE = kap[c] * th + kap[w] * (1.0 - th);
        double val = A.dt * (-A.sigma * kE) * (H[c] - H[w]);
        val *= iRho / G.dx;
        val += us[c] - iRho * (ps[c] - ps[w]) / G.dx;
        u[c] = val;
    }
    if (j >= G.g + 1) {
        const int64_t s = G.idx(i, j - 1);
        double thH, th;
        face_theta_w(ls[s], ls[c], thH, th);
        const double iRho = 1.0 / (A.rho_h * thH + A.rho_l * (1.0 - thH));
        const double kE = kap[c] * th + kap[s] * (1.0 - th);
        double val = A.dt * (-A.sigma * kE) * (H[c] - H[s]);
        val *= iRho / G.dy;
        val += vs[c] - iRho * (ps[c] - ps[s]) / G.dy;
        v[c] = val;
    }
}

// ------------------------------------------------------------------ CFL
// max over the interior of |u_{i+1} u_i / 2| and |v_{j+1} v_j / 2| (src/mac2d.cpp:1858-1862):
// block maxima into partials, then one block reduces them (max is order-free, so the result
// does not depend on the split)
// (kap given: also the max of |kappa|, for the axisymmetric time step, src/mac2dAxisym.cpp:2103)
__global__ __launch_bounds__(256) void k_velmax(Geom G, const double* __restrict__ u, const double* __restrict__ v,
                                                const double* __restrict__ kap, double* __restrict__ partials) {
    double um = 0.0, vm = 0.0, km = 0.0;
    for (int i = G.g + blockIdx.y; i < G.nx + G.g; i += gridDim.y)
        for (int j = G.g + blockIdx.x * 256 + threadIdx.x; j < G.ny + G.g; j += gridDim.x * 256) {
            um = smax(um, fabs((u[G.idx(i + 1, j)] * u[G.idx(i, j)]) * 0.5));
            vm = smax(vm, fabs((v[G.idx(i, j + 1)] * v[G.idx(i, j)]) * 0.5));
            if (kap) km = smax(km, fabs(kap[G.idx(i, j)]));
        }
    __shared__ double su[256], sv[256], sk[256];
    su[threadIdx.x] = um;
    sv[threadIdx.x] = vm;
    sk[threadIdx.x] = km;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            su[threadIdx.x] = smax(su[threadIdx.x], su[threadIdx.x + s]);
            sv[threadIdx.x] = smax(sv[threadIdx.x], sv[threadIdx.x + s]);
            sk[threadIdx.x] = smax(sk[threadIdx.x], sk[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int b = blockIdx.y * gridDim.x + blockIdx.x, nb = gridDim.x * gridDim.y;
        partials[b] = su[0];
        partials[nb + b] = sv[0];
        partials[2 * nb + b] = sk[0];
    }
}

__global__ __launch_bounds__(256) void k_velmax_final(const double* __restrict__ partials, int nb,
                                                      double* __restrict__ out) {
    double um = 0.0, vm = 0.0, km = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        um = smax(um, partials[b]);
        vm = smax(vm, partials[nb + b]);
        km = smax(km, partials[2 * nb + b]);
    }
    __shared__ double su[256], sv[256], sk[256];
    su[threadIdx.x] = um;
    sv[threadIdx.x] = vm;
    sk[threadIdx.x] = km;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            su[threadIdx.x] = smax(su[threadIdx.x], su[threadIdx.x + s]);
            sv[threadIdx.x] = smax(sv[threadIdx.x], sv[threadIdx.x + s]);
            sk[threadIdx.x] = smax(sk[threadIdx.x], sk[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = su[0];
        out[1] = sv[0];
        out[2] = sk[0];
    }
}

// ------------------------------------------------------------------ momentum predictor
// mu = mu_L + (mu_H - mu_L) H on the interior, 0 on ghosts (src/mac2d.cpp:606-608)
__device__ __forceinline__ double mu_at(const Geom& G, const double* Hs, double muH, double muL, int i, int j) {
#pragma clang fp contract(off)
    if (i < G.g || i >= G.nx + G.g || j < G.g || j >= G.ny + G.g) return 0.0;
    return muL + (muH - muL) * Hs[G.idx(i, j)];
}

struct MomArgs {
    double rho_h, rho_l, mu_h, mu_l, dt;
    double irh, irl, irlrh;  // 1 / rho_h, 1 / rho_l, 1 / (rho_l rho_h), rounded once on the host
    double gu, gv;    // gravity source on U faces / V faces
    int gv_from_i;    // V gravity only where i >= g + 1 (the U-grid array of mac2d.cpp:294)
    int viscous;      // kRe > 0
    int mode;         // 0: s + dt F; 1: 0.75 u0 + 0.25 (s + dt F); 2: u0/3 + 2/3 (s + dt F); 3: u0/2 + (s + dt F)/2
                      // 4: F; 5: convection term; 6: viscous term
};

// 1 / rho at a U face (harmonic mean across the interface, src/mac2d.cpp:570-809); the reciprocals
// of the densities are the host's correctly rounded 1 / rho_h, 1 / rho_l and 1 / (rho_l rho_h), the
// values the reference's expressions produce
__device__ __forceinline__ double irho_harmonic(const MomArgs& A, double a, double b) {
#pragma clang fp contract(off)
    if (a >= 0 && b >= 0) return A.irh;
    else if (a <= 0 && b <= 0) return A.irl;
    else if (a > 0 && b <= 0) {
        const double th = fabs(a) / (fabs(a) + fabs(b));
        return A.irlrh / (A.irl * th + A.irh * (1.0 - th));
    } else {
        const double th = fabs(a) / (fabs(a) + fabs(b));
        return A.irlrh / (A.irh * th + A.irl * (1.0 - th));
    }
}

// 1 / rho at a V face (arithmetic mean across the interface, the FV override of
// src/mac2d.cpp:1000-1001)
__device__ __forceinline__ double irho_arith(const MomArgs& A, double a, double b) {
#pragma clang fp contract(off)
    const double rH = A.rho_h, rL = A.rho_l;
    if (a >= 0 && b >= 0) return A.irh;
    else if (a <= 0 && b <= 0) return A.irl;
    else if (a > 0 && b <= 0) {
        const double th = fabs(a) / (fabs(a) + fabs(b));
        return 1.0 / (rH * th + rL * (1.0 - th));
    } else {
        const double th = fabs(a) / (fabs(a) + fabs(b));
        return 1.0 / (rL * th + rH * (1.0 - th));
    }
}

// -cU + vU + gU at a U face and the RK combination (src/mac2d.cpp:254-279, 308-387, 570-809)
// Both momentum kernels run on MJ x MI face tiles (j fastest, 4 rows per thread).  A block stages
// its tile of u and v with the halos the stencils read, and mu and the level set at the cells
// around its faces, in LDS, fetching every one of them before the first LDS store; the WENO fluxes
// are then formed from the staged values (the product flux once per face into an LDS tile with a
// 3-face halo, the squared flux where the WENO stencil reads it).
constexpr int MJ = 64, MI = 16;

// mu at the staged cells: mu_at over Hs values already in registers (FULL: over the whole padded
// array, as the axisymmetric solver computes it, src/mac2dAxisym.cpp:607-609)
template <bool FULL, int TH, int TW, int NT>
__device__ __forceinline__ void mu_store(const Geom& G, const TileRegs<TH, TW, NT>& t, double (*s)[TW], int ib, int jb,
                                         double muH, double muL) {
#pragma unroll
    for (int k = 0; k < TileRegs<TH, TW, NT>::NL; k++) {
        const int e = static_cast<int>(threadIdx.x) + k * NT;
        if ((TH * TW) % NT == 0 || e < TH * TW) {
            const int r = e / TW, c = e - r * TW;
            const int i = ib + r, j = jb + c;
            double m;
            {
#pragma clang fp contract(off)
                m = (!FULL && (i < G.g || i >= G.nx + G.g || j < G.g || j >= G.ny + G.g)) ? 0.0
                                                                                      : muL + (muH - muL) * t.v[k];
            }
            s[r][c] = m;
        }
    }
}

// AXI: the axisymmetric solver's U right-hand side (GetRHSU, src/mac2dAxisym.cpp:264-292): the same
// convection, the viscous term of AddExternalViscosityFU (:590-753) and u + dt (-cU + gU + vU)
template <bool POW2, bool AXI>
__global__ __launch_bounds__(256) void k_mom_u(Geom G, MomArgs A, const double* __restrict__ u,
                                               const double* __restrict__ v, const double* __restrict__ ls,
                                               const double* __restrict__ Hs, const double* __restrict__ u0,
                                               double* __restrict__ out) {
    constexpr int UH = MI + 6, UW = MJ + 6;  // u at (i0 - 3 + a, j0 - 3 + b)
    constexpr int VH = MI + 1, VW = MJ + 7;  // v at (i0 - 1 + a, j0 - 3 + b)
    constexpr int CH = MI + 1, CW = MJ + 2;  // mu, level set at cells (i0 - 1 + a, j0 - 1 + b)
    __shared__ double sU[UH][UW];
    __shared__ double sV[VH][VW];
    __shared__ double sFy[MI][MJ + 6];  // u vbar along j
    __shared__ double sMu[CH][CW];
    __shared__ double sLs[CH][CW];
    const int tj = threadIdx.x & (MJ - 1), ti = threadIdx.x / MJ;
    const int j0 = G.g + blockIdx.x * MJ, i0 = G.g + 1 + blockIdx.y * MI;
    const int ipad = G.nx + 2 * G.g, jpad = G.ny + 2 * G.g;
    const int iend = G.nx + G.g, jend = G.ny + G.g;
    TileRegs<UH, UW, 256> tu;
    TileRegs<VH, VW, 256> tv;
    TileRegs<CH, CW, 256> th, tl;
    tile_fetch(G, u, tu, i0 - 3, j0 - 3);
    tile_fetch(G, v, tv, i0 - 1, j0 - 3);
    if (A.viscous) {
        tile_fetch(G, Hs, th, i0 - 1, j0 - 1);
        tile_fetch(G, ls, tl, i0 - 1, j0 - 1);
    }
    constexpr int NR = MI / (256 / MJ);
    double u0R[NR];
    const bool rk = A.mode >= 1 && A.mode <= 3;
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int i = i0 + ti + (256 / MJ) * k, j = j0 + tj;
        const bool ok = rk && i < iend && j < jend;
        const double x = u0[ok ? G.idx(i, j) : 0];
        u0R[k] = ok ? x : 0.0;
    }
    tile_store(tu, sU);
    tile_store(tv, sV);
    if (A.viscous) {
        mu_store<AXI>(G, th, sMu, i0 - 1, j0 - 1, A.mu_h, A.mu_l);
        tile_store(tl, sLs);
    }
    __syncthreads();
    // tmpV of AddConvectionFU at a U face (src/mac2d.cpp:314-316) from the staged v, 0 elsewhere
    auto vbar = [&](int i, int j) -> double {
        if (j < G.g || j >= G.ny + G.g || i < G.g + 1 || i >= G.nx + G.g) return 0.0;
        const int a = i - (i0 - 1), b = j - (j0 - 3);
        return (sV[a][b] + sV[a - 1][b] + sV[a - 1][b + 1] + sV[a][b + 1]) * 0.25;
    };
    for (int e = threadIdx.x; e < MI * (MJ + 6); e += 256) {
        const int r = e / (MJ + 6), cc = e - r * (MJ + 6);
        const int i = i0 + r, j = j0 - 3 + cc;
        sFy[r][cc] = (i < ipad && j < jpad) ? sU[r + 3][cc] * vbar(i, j) : 0.0;
    }
    __syncthreads();
    const double idx = 1.0 / G.dx, idy = 1.0 / G.dy, epsx = 1e-6 * G.dx * G.dx, epsy = 1e-6 * G.dy * G.dy;
    for (int k = 0; k < NR; k++) {
    const int r = ti + (256 / MJ) * k;
    const int i = i0 + r, j = j0 + tj;
    if (i >= iend || j >= jend) continue;
    const int64_t c = G.idx(i, j);
    const int R = r + 3, C = tj + 3;  // the face in sU
    // convection: WENO of u^2 along i and of u * vbar along j
    double lxp, lxm, lyp, lym;
    {
        double f[7];
#pragma unroll
        for (int q = 0; q < 7; q++) {
            const double w = sU[R - 3 + q][C];
            f[q] = w * w;
        }
        weno5(f[1] - f[0], f[2] - f[1], f[3] - f[2], f[4] - f[3], f[5] - f[4], f[6] - f[5], idx, epsx, lxp,
              lxm);
    }
    weno5(sFy[r][tj + 1] - sFy[r][tj], sFy[r][tj + 2] - sFy[r][tj + 1], sFy[r][tj + 3] - sFy[r][tj + 2],
          sFy[r][tj + 4] - sFy[r][tj + 3], sFy[r][tj + 5] - sFy[r][tj + 4], sFy[r][tj + 6] - sFy[r][tj + 5],
          idy, epsy, lyp, lym);
    const double aX = sU[R][C], aY = vbar(i, j);
    const double cU = 0.5 * (lxp + lxm) + 0.5 * (lyp + lym) - aX * (0.5 * (lxp - lxm)) - aY * (0.5 * (lyp - lym));
    double dU = 0.0;
    if (AXI && A.viscous) {
#pragma clang fp contract(off)
        // d/dz (mu dv/dr) over the arithmetic-mean density of the S-N half points
        const int a = i - (i0 - 1), b = j - (j0 - 3);
        const double vW = sV[a - 1][b], vW_N = sV[a - 1][b + 1], vM = sV[a][b], vN = sV[a][b + 1];
        const int a0 = r + 1, b0 = tj + 1;
#define LS(di, dj) sLs[a0 + (di)][b0 + (dj)]
#define MU(di, dj) sMu[a0 + (di)][b0 + (dj)]
        const double lsUNHalf = 0.25 * (LS(0, 0) + LS(-1, 0) + LS(-1, 1) + LS(0, 1));
        const double lsUSHalf = 0.25 * (LS(0, 0) + LS(-1, 0) + LS(-1, -1) + LS(0, -1));
        const double muUNHalf = 0.25 * (MU(0, 0) + MU(-1, 0) + MU(-1, 1) + MU(0, 1));
        const double muUSHalf = 0.25 * (MU(0, 0) + MU(-1, 0) + MU(-1, -1) + MU(0, -1));
#undef LS
#undef MU
        const double rhoSN = rho_eff_axi(A.rho_h, A.rho_l, lsUSHalf, lsUNHalf);
        const double visZ = wdiv<POW2>(wdiv<POW2>(muUNHalf * (vN - vW_N), G.dx, idx)
                                           - wdiv<POW2>(muUSHalf * (vM - vW), G.dx, idx), G.dy, idy);
        dU = visZ / rhoSN;
    } else if (!AXI && A.viscous) {
#pragma clang fp contract(off)
        const double dx = G.dx, dy = G.dy;
        const double uM = sU[R][C], uW = sU[R - 1][C], uE = sU[R + 1][C];
        const double uS = sU[R][C - 1], uN = sU[R][C + 1];
        const int a = i - (i0 - 1), b = j - (j0 - 3);  // the face's own cell in sV
        const double vW = sV[a - 1][b], vW_N = sV[a - 1][b + 1], vM = sV[a][b], vN = sV[a][b + 1];
        // tile cell (i + di, j + dj) of the staged mu / level set
        const int a0 = r + 1, b0 = tj + 1;
#define LS(di, dj) sLs[a0 + (di)][b0 + (dj)]
        const double lsW = LS(-1, 0), lsM = LS(0, 0);
        const double lsUNHalf = 0.25 * (LS(0, 0) + LS(-1, 0) + LS(-1, 1) + LS(0, 1));
        const double lsUSHalf = 0.25 * (LS(0, 0) + LS(-1, 0) + LS(-1, -1) + LS(0, -1));
#undef LS
        const double iWE = irho_harmonic(A, lsW, lsM);
        const double iSN = irho_harmonic(A, lsUSHalf, lsUNHalf);
#define MU(a, b) sMu[a0 + (a) - i][b0 + (b) - j]
        const double muU_X_W = wdiv<POW2>(MU(i - 1, j) * (uM - uW), dx, idx);
        const double muU_X_E = wdiv<POW2>(MU(i, j) * (uE - uM), dx, idx);
        const double muU_Y_S =
            wdiv<POW2>(0.25 * (MU(i, j) + MU(i - 1, j) + MU(i - 1, j - 1) + MU(i, j - 1)) * (uM - uS), dy, idy);
        const double muU_Y_N =
            wdiv<POW2>(0.25 * (MU(i, j) + MU(i - 1, j) + MU(i - 1, j + 1) + MU(i, j + 1)) * (uN - uM), dy, idy);
        const double muV_X_S =
            wdiv<POW2>(0.25 * (MU(i, j) + MU(i - 1, j) + MU(i - 1, j - 1) + MU(i, j - 1)) * (vM - vW), dx, idx);
        const double muV_X_N =
            wdiv<POW2>(0.25 * (MU(i, j) + MU(i - 1, j) + MU(i - 1, j + 1) + MU(i, j + 1)) * (vN - vW_N), dx, idx);
#undef MU
        const double visX = wdiv<POW2>(2.0 * (muU_X_E - muU_X_W), dx, idx);
        const double visY = wdiv<POW2>(muU_Y_N - muU_Y_S, dy, idy) + wdiv<POW2>(muV_X_N - muV_X_S, dy, idy);
        dU = visX * iWE + visY * iSN;
    }
    const double F = AXI ? (-cU + A.gu) + dU : -cU + dU + A.gu;
    const double t = sU[R][C] + A.dt * F;
    double o;
    if (A.mode == 0) o = t;
    else if (A.mode == 1) o = 0.75 * u0R[k] + 0.25 * t;
    else if (A.mode == 2) o = 1.0 / 3.0 * u0R[k] + 2.0 / 3.0 * t;
    else if (A.mode == 3) o = 0.5 * u0R[k] + 0.5 * t;
    else if (A.mode == 4) o = F;
    else if (A.mode == 5) o = cU;
    else o = dU;
    out[c] = o;
    }
}

// -cV + vV + gV at a V face and the RK combination (src/mac2d.cpp:281-306, 389-463, 811-1041)
// AXI: the axisymmetric solver's V right-hand side (GetRHSV, src/mac2dAxisym.cpp:294-322): the same
// convection, the viscous term of AddExternalViscosityFV (:755-910, with the new U-face velocity
// uh) and v + dt (-cV + gV + vV)
template <bool POW2, bool AXI>
__global__ __launch_bounds__(256) void k_mom_v(Geom G, MomArgs A, const double* __restrict__ u,
                                               const double* __restrict__ v, const double* __restrict__ ls,
                                               const double* __restrict__ Hs, const double* __restrict__ v0,
                                               const double* __restrict__ uh, double* __restrict__ out) {
    constexpr int VH = MI + 6, VW = MJ + 6;  // v at (i0 - 3 + a, j0 - 3 + b)
    constexpr int UH = MI + 7, UW = MJ + 1;  // u at (i0 - 3 + a, j0 - 1 + b)
    constexpr int CH = MI + 2, CW = MJ + 1;  // mu, level set at cells (i0 - 1 + a, j0 - 1 + b)
    __shared__ double sV[VH][VW];
    __shared__ double sU[UH][UW];
    __shared__ double sFx[MI + 6][MJ];  // v ubar along i
    __shared__ double sMu[CH][CW];
    __shared__ double sLs[CH][CW];
    constexpr int HH = AXI ? MI + 1 : 1, HW = AXI ? MJ + 1 : 1;  // uh at (i0 + a, j0 - 1 + b)
    __shared__ double sUh[HH][HW];
    const int tj = threadIdx.x & (MJ - 1), ti = threadIdx.x / MJ;
    const int j0 = G.g + 1 + blockIdx.x * MJ, i0 = G.g + blockIdx.y * MI;
    const int ipad = G.nx + 2 * G.g, jpad = G.ny + 2 * G.g;
    const int iend = G.nx + G.g, jend = G.ny + G.g;
    TileRegs<VH, VW, 256> tv;
    TileRegs<UH, UW, 256> tu;
    TileRegs<CH, CW, 256> th, tl;
    TileRegs<HH, HW, 256> tuh;
    tile_fetch(G, v, tv, i0 - 3, j0 - 3);
    tile_fetch(G, u, tu, i0 - 3, j0 - 1);
    if (A.viscous) {
        tile_fetch(G, Hs, th, i0 - 1, j0 - 1);
        tile_fetch(G, ls, tl, i0 - 1, j0 - 1);
        if (AXI) tile_fetch(G, uh, tuh, i0, j0 - 1);
    }
    constexpr int NR = MI / (256 / MJ);
    double v0R[NR];
    const bool rk = A.mode >= 1 && A.mode <= 3;
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int i = i0 + ti + (256 / MJ) * k, j = j0 + tj;
        const bool ok = rk && i < iend && j < jend;
        const double x = v0[ok ? G.idx(i, j) : 0];
        v0R[k] = ok ? x : 0.0;
    }
    tile_store(tv, sV);
    tile_store(tu, sU);
    if (A.viscous) {
        mu_store<AXI>(G, th, sMu, i0 - 1, j0 - 1, A.mu_h, A.mu_l);
        tile_store(tl, sLs);
        if (AXI) tile_store(tuh, sUh);
    }
    __syncthreads();
    // tmpU of AddConvectionFV at a V face (src/mac2d.cpp:395-397) from the staged u, 0 elsewhere
    auto ubar = [&](int i, int j) -> double {
        if (j < G.g + 1 || j >= G.ny + G.g || i < G.g || i >= G.nx + G.g) return 0.0;
        const int a = i - (i0 - 3), b = j - (j0 - 1);
        return (sU[a][b] + sU[a][b - 1] + sU[a + 1][b - 1] + sU[a + 1][b]) * 0.25;
    };
    for (int r = ti; r < MI + 6; r += 256 / MJ) {
        const int i = i0 - 3 + r, j = j0 + tj;
        sFx[r][tj] = (i < ipad && j < jpad) ? sV[r][tj + 3] * ubar(i, j) : 0.0;
    }
    __syncthreads();
    const double idx = 1.0 / G.dx, idy = 1.0 / G.dy, epsx = 1e-6 * G.dx * G.dx, epsy = 1e-6 * G.dy * G.dy;
    for (int k = 0; k < NR; k++) {
    const int r = ti + (256 / MJ) * k;
    const int i = i0 + r, j = j0 + tj;
    if (i >= iend || j >= jend) continue;
    const int64_t c = G.idx(i, j);
    const int R = r + 3, C = tj + 3;  // the face in sV
    double lxp, lxm, lyp, lym;
    weno5(sFx[r + 1][tj] - sFx[r][tj], sFx[r + 2][tj] - sFx[r + 1][tj], sFx[r + 3][tj] - sFx[r + 2][tj],
          sFx[r + 4][tj] - sFx[r + 3][tj], sFx[r + 5][tj] - sFx[r + 4][tj], sFx[r + 6][tj] - sFx[r + 5][tj],
          idx, epsx, lxp, lxm);
    {
        double f[7];
#pragma unroll
        for (int q = 0; q < 7; q++) {
            const double w = sV[R][C - 3 + q];
            f[q] = w * w;
        }
        weno5(f[1] - f[0], f[2] - f[1], f[3] - f[2], f[4] - f[3], f[5] - f[4], f[6] - f[5], idy, epsy, lyp,
              lym);
    }
    const double aX = ubar(i, j), aY = sV[R][C];
    const double cV = 0.5 * (lxp + lxm) + 0.5 * (lyp + lym) - aX * (0.5 * (lxp - lxm)) - aY * (0.5 * (lyp - lym));
    double dV = 0.0;
    if (AXI && A.viscous) {
#pragma clang fp contract(off)
        // d/dr (mu du/dz) over the W-E half points' density, plus mu / r d(u*)/dz over the S-N
        // cells' density; r without the base radius, as the reference computes it
        const int a = i - (i0 - 3), b = j - (j0 - 1);
        const double uM = sU[a][b], uE = sU[a + 1][b], uS = sU[a][b - 1], uS_E = sU[a + 1][b - 1];
        const int a0 = r + 1, b0 = tj + 1;
#define LS(di, dj) sLs[a0 + (di)][b0 + (dj)]
#define MU(di, dj) sMu[a0 + (di)][b0 + (dj)]
        const double lsVWHalf = 0.25 * (LS(0, 0) + LS(-1, 0) + LS(-1, -1) + LS(0, -1));
        const double lsVEHalf = 0.25 * (LS(0, 0) + LS(1, 0) + LS(1, -1) + LS(0, -1));
        const double muVWHalf = 0.25 * (MU(0, 0) + MU(-1, 0) + MU(-1, -1) + MU(0, -1));
        const double muVEHalf = 0.25 * (MU(0, 0) + MU(1, 0) + MU(1, -1) + MU(0, -1));
        const double rhoWE = rho_eff_axi(A.rho_h, A.rho_l, lsVWHalf, lsVEHalf);
        const double rhoSN = rho_eff_axi(A.rho_h, A.rho_l, LS(0, -1), LS(0, 0));
        const double muSN = 0.5 * (MU(0, 0) + MU(0, -1));
#undef LS
#undef MU
        const double rPM = (i + 0.5 - G.g) * G.dx;
        const double visR = wdiv<POW2>(wdiv<POW2>(muVEHalf * (uE - uS_E), G.dy, idy)
                                           - wdiv<POW2>(muVWHalf * (uM - uS), G.dy, idy), G.dx, idx);
        const int ha = r, hb = tj + 1;  // the face's own U face (i, j) in sUh
        const double duh = 0.5 * (sUh[ha][hb] + sUh[ha + 1][hb]) - 0.5 * (sUh[ha][hb - 1] + sUh[ha + 1][hb - 1]);
        const double visZ = wdiv<POW2>(muSN / rPM * duh, G.dy, idy);
        dV = visR / rhoWE + visZ / rhoSN;
    } else if (!AXI && A.viscous) {
#pragma clang fp contract(off)
        const double dx = G.dx, dy = G.dy;
        const double vW = sV[R - 1][C], vE = sV[R + 1][C], vS = sV[R][C - 1];
        const double vN = sV[R][C + 1], vM = sV[R][C];
        const int a = i - (i0 - 3), b = j - (j0 - 1);  // the face's own cell in sU
        const double uM = sU[a][b], uE = sU[a + 1][b], uS = sU[a][b - 1], uE_S = sU[a + 1][b - 1];
        // tile cell (i + di, j + dj) of the staged mu / level set
        const int a0 = r + 1, b0 = tj + 1;
#define LS(di, dj) sLs[a0 + (di)][b0 + (dj)]
        const double lsVWHalf = 0.25 * (LS(0, 0) + LS(-1, 0) + LS(-1, -1) + LS(0, -1));
        const double lsVEHalf = 0.25 * (LS(0, 0) + LS(1, 0) + LS(1, -1) + LS(0, -1));
        const double iWE = irho_arith(A, lsVWHalf, lsVEHalf);
        const double iSN = irho_arith(A, LS(0, -1), LS(0, 0));
#undef LS
#define MU(a, b) sMu[a0 + (a) - i][b0 + (b) - j]
        const double muU_Y_W =
            wdiv<POW2>(0.25 * (MU(i, j) + MU(i - 1, j) + MU(i - 1, j - 1) + MU(i, j - 1)) * (uM - uS), dy, idy);
        const double muU_Y_E =
            wdiv<POW2>(0.25 * (MU(i, j) + MU(i + 1, j) + MU(i + 1, j - 1) + MU(i, j - 1)) * (uE - uE_S), dy, idy);
        const double muV_X_W =
            wdiv<POW2>(0.25 * (MU(i, j) + MU(i - 1, j) + MU(i - 1, j - 1) + MU(i, j - 1)) * (vM - vW), dx, idx);
        const double muV_X_E =
            wdiv<POW2>(0.25 * (MU(i, j) + MU(i + 1, j) + MU(i + 1, j - 1) + MU(i, j - 1)) * (vE - vM), dx, idx);
        const double muV_Y_S = wdiv<POW2>(MU(i, j - 1) * (vM - vS), dy, idy);
        const double muV_Y_N = wdiv<POW2>(MU(i, j) * (vN - vM), dy, idy);
#undef MU
        const double visX = wdiv<POW2>(muU_Y_E - muU_Y_W, dx, idx) + wdiv<POW2>(muV_X_E - muV_X_W, dx, idx);
        const double visY = wdiv<POW2>(2.0 * (muV_Y_N - muV_Y_S), dy, idy);
        dV = visX * iWE + visY * iSN;
    }
    const double gv = (A.gv_from_i && i < G.g + 1) ? 0.0 : A.gv;
    const double F = AXI ? (-cV + gv) + dV : -cV + dV + gv;
    const double t = sV[R][C] + A.dt * F;
    double o;
    if (A.mode == 0) o = t;
    else if (A.mode == 1) o = 0.75 * v0R[k] + 0.25 * t;
    else if (A.mode == 2) o = 1.0 / 3.0 * v0R[k] + 2.0 / 3.0 * t;
    else if (A.mode == 3) o = 0.5 * v0R[k] + 0.5 * t;
    else if (A.mode == 4) o = F;
    else if (A.mode == 5) o = cV;
    else o = dV;
    out[c] = o;
    }
}

}  // namespace asd2

using namespace asd2;

namespace asd2 {
namespace {
// h = {max |u_{i+1} u_i / 2|, max |v_{j+1} v_j / 2|, max |kappa| (0 without kappa)} over the interior
int velocity_maxima(asd2_ctx* ctx, const double* d_u, const double* d_v, const double* d_kappa, double h[3]) {
    const Geom& G = ctx->geo;
    double *partials, *out;
    const dim3 grid(std::max(1, std::min(8, (G.ny + 255) / 256)), std::max(1, std::min(128, G.nx)));
    const int nb = static_cast<int>(grid.x * grid.y);
    ASD2_TRY(ws(ctx, "dt_partials", 3 * nb, &partials));
    ASD2_TRY(ws(ctx, "dt_out", 3, &out));
    hipLaunchKernelGGL(k_velmax, grid, dim3(256), 0, ctx->stream, G, d_u, d_v, d_kappa, partials);
    hipLaunchKernelGGL(k_velmax_final, dim3(1), dim3(256), 0, ctx->stream, partials, nb, out);
    ASD2_LAUNCH_CHECK();
    ASD2_HIP(hipMemcpyAsync(h, out, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ASD2_HIP(hipStreamSynchronize(ctx->stream));
    return ASD2_OK;
}
// momentum arguments of a fluid: densities, their reciprocals rounded once on the host, gravity on
// the U and V faces (the V equation takes the U-grid array unless gravity_fix, src/mac2d.cpp:294,
// src/mac2dAxisym.cpp:312), kRe > 0 for the viscous terms
MomArgs mom_args(const asd2_fluid* fl, int mode) {
    MomArgs A{};
    A.rho_h = fl->rho_h; A.rho_l = fl->rho_l; A.mu_h = fl->mu_h; A.mu_l = fl->mu_l; A.dt = fl->dt;
    A.irh = 1.0 / fl->rho_h; A.irl = 1.0 / fl->rho_l; A.irlrh = 1.0 / (fl->rho_l * fl->rho_h);
    const bool fr_ok = (fl->fr != 0 || fl->g != 0.0) && !std::isnan(fl->fr) && !std::isinf(fl->fr);
    A.gu = (fr_ok && fl->g_axis == 0) ? -fl->g : 0.0;
    if (fl->gravity_fix) { A.gv = (fr_ok && fl->g_axis == 1) ? -fl->g : 0.0; A.gv_from_i = 0; }
    else { A.gv = A.gu; A.gv_from_i = 1; }
    A.viscous = fl->re > 0.0 ? 1 : 0;
    A.mode = mode;
    return A;
}

// one momentum launch on the interior U (grid_kind = ASD2_GRID_U) or V faces; s0 is the RK stage's
// u0 / v0, uh the new U-face velocity of the axisymmetric V term
int launch_mom(asd2_ctx* ctx, const MomArgs& A, int grid_kind, bool axi, const double* us, const double* vs,
               const double* ls, const double* Hs, const double* s0, const double* uh, double* out) {
    const Geom& G = ctx->geo;
    const bool pow2 = is_pow2(G.dx) && is_pow2(G.dy);  // divisions by dx, dy are exact products
    hipStream_t st = ctx->stream;
    prof_begin(ctx, ASD2_KID_MOMENTUM);
    if (grid_kind == ASD2_GRID_U) {
        const dim3 gr(cdiv(G.ny, MJ), cdiv(G.nx - 1, MI));
        if (axi && pow2) hipLaunchKernelGGL((k_mom_u<true, true>), gr, dim3(256), 0, st, G, A, us, vs, ls, Hs, s0, out);
        else if (axi) hipLaunchKernelGGL((k_mom_u<false, true>), gr, dim3(256), 0, st, G, A, us, vs, ls, Hs, s0, out);
        else if (pow2) hipLaunchKernelGGL((k_mom_u<true, false>), gr, dim3(256), 0, st, G, A, us, vs, ls, Hs, s0, out);
        else hipLaunchKernelGGL((k_mom_u<false, false>), gr, dim3(256), 0, st, G, A, us, vs, ls, Hs, s0, out);
    } else {
        const dim3 gr(cdiv(G.ny - 1, MJ), cdiv(G.nx, MI));
        if (axi && pow2)
            hipLaunchKernelGGL((k_mom_v<true, true>), gr, dim3(256), 0, st, G, A, us, vs, ls, Hs, s0, uh, out);
        else if (axi)
            hipLaunchKernelGGL((k_mom_v<false, true>), gr, dim3(256), 0, st, G, A, us, vs, ls, Hs, s0, uh, out);
        else if (pow2)
            hipLaunchKernelGGL((k_mom_v<true, false>), gr, dim3(256), 0, st, G, A, us, vs, ls, Hs, s0, uh, out);
        else hipLaunchKernelGGL((k_mom_v<false, false>), gr, dim3(256), 0, st, G, A, us, vs, ls, Hs, s0, uh, out);
    }
    prof_end(ctx, ASD2_KID_MOMENTUM);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

}  // namespace
}  // namespace asd2

extern "C" {

int asd2_mac_kappa(asd2_ctx* ctx, const double* d_ls, double* d_kappa) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_kappa) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    hipLaunchKernelGGL(k_kappa, dim3(cdiv(G.ny, 64), cdiv(G.nx, 4)), dim3(256), 0, ctx->stream, G, d_ls, d_kappa);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_mac_heaviside(asd2_ctx* ctx, const double* d_ls, double* d_H, int smooth) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_H) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    const double eps = std::min(G.dx, G.dy) * 1.5;
    hipLaunchKernelGGL(k_heaviside, dim3(2048), dim3(256), 0, ctx->stream, G.size(), eps, smooth ? 1 : 0, d_ls, d_H);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_mac_divergence(asd2_ctx* ctx, const double* d_u, const double* d_v, double* d_div) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_u || !d_v || !d_div) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    hipLaunchKernelGGL(k_divergence, rows_grid(G), dim3(256), 0, ctx->stream, G, d_u, d_v, d_div);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_mac_pressure(asd2_ctx* ctx, const double* d_ps, double* d_p, double dt) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ps || !d_p) { set_error("null array"); return ASD2_ERR_ARG; }
    if (!(dt != 0.0)) { set_error("dt must be nonzero"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    hipLaunchKernelGGL(k_pressure, rows_grid(G), dim3(256), 0, ctx->stream, G, d_ps, d_p, dt);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_mac_update_vel(asd2_ctx* ctx, const asd2_fluid* fl, double* d_u, double* d_v, const double* d_us,
                        const double* d_vs, const double* d_ps, const double* d_ls, const double* d_kappa,
                        const double* d_H) {
    ASD2_TRY(check_ctx(ctx));
    if (!fl || !d_u || !d_v || !d_us || !d_vs || !d_ps || !d_ls || !d_kappa || !d_H) {
        set_error("null argument");
        return ASD2_ERR_ARG;
    }
    const Geom& G = ctx->geo;
    ProjArgs A{fl->rho_h, fl->rho_l, fl->sigma, fl->dt};
    hipLaunchKernelGGL(k_update_vel, dim3(cdiv(G.ny, 64), cdiv(G.nx, 4)), dim3(256), 0, ctx->stream, G, A, d_u, d_v,
                       d_us, d_vs, d_ps, d_ls, d_kappa, d_H);
    ASD2_LAUNCH_CHECK();
    ASD2_TRY(bc_apply_uv(ctx, ASD2_GRID_U, d_u));
    ASD2_TRY(bc_apply_uv(ctx, ASD2_GRID_V, d_v));
    return ASD2_OK;
}


// MACSolver2DAxisym::UpdateDt (src/mac2dAxisym.cpp:2094-2128): the capillary limit from max |kappa|,
// no viscous term in the bound
int asd2_mac_update_dt_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_u, const double* d_v,
                              const double* d_kappa, double* dt_out) {
    ASD2_TRY(check_ctx(ctx));
    if (!fl || !d_u || !d_v || !d_kappa || !dt_out) { set_error("null argument"); return ASD2_ERR_ARG; }
    double h[3];
    ASD2_TRY(velocity_maxima(ctx, d_u, d_v, d_kappa, h));
    const double dr = ctx->geo.dx, dz = ctx->geo.dy;
    const double C = h[0] / dr + h[1] / dz;
    const double Gf = std::max(std::sqrt(std::fabs(fl->g) / dz), std::sqrt(std::fabs(fl->g) / dr));
    const double S = std::sqrt((fl->sigma * h[2]) / (std::min(fl->rho_h, fl->rho_l) * std::pow(std::min(dr, dz), 2.0)));
    double dt = std::numeric_limits<double>::max();
    dt = std::min(dt, fl->cfl / (0.5 * (C + std::sqrt(std::pow(C, 2.0) + 4.0 * std::pow(Gf, 2.0) +
                                                     4.0 * std::pow(S, 2.0)))));
    *dt_out = dt;
    if (!std::isfinite(dt)) { set_error("dt nan/inf"); return ASD2_ERR_NONFINITE; }
    return ASD2_OK;
}

int asd2_mac_update_dt(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_u, const double* d_v, double* dt_out) {
    ASD2_TRY(check_ctx(ctx));
    if (!fl || !d_u || !d_v || !dt_out) { set_error("null argument"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    double h[3];
    ASD2_TRY(velocity_maxima(ctx, d_u, d_v, nullptr, h));
    const double dx = G.dx, dy = G.dy;
    const double C = h[0] / dx + h[1] / dy;
    const double V = std::max(fl->mu_h, fl->mu_l) * (2.0 / (dx * dx + dy * dy));
    const double Gf = std::max(std::sqrt(std::fabs(fl->g) / dy), std::sqrt(std::fabs(fl->g) / dx));
    const double S = std::sqrt((fl->sigma / std::min(dx, dy)) /
                               (std::min(fl->rho_h, fl->rho_l) * std::pow(std::min(dx, dy), 2.0)));
    double dt = std::numeric_limits<double>::max();
    dt = std::min(dt, fl->cfl / (0.5 * (C + V + std::sqrt(std::pow(C + V, 2.0) + 4.0 * Gf * Gf + 4.0 * S * S))));
    *dt_out = dt;
    if (!std::isfinite(dt)) { set_error("dt nan/inf"); return ASD2_ERR_NONFINITE; }
    return ASD2_OK;
}

int asd2_mac_intermediate_vel(asd2_ctx* ctx, const asd2_fluid* fl, int time_order, const double* d_ls,
                              const double* d_u, const double* d_v, double* d_uhat, double* d_vhat,
                              const double* d_Hs) {
    ASD2_TRY(check_ctx(ctx));
    if (!fl || !d_ls || !d_u || !d_v || !d_uhat || !d_vhat || !d_Hs || time_order < 0 || time_order > 2) {
        set_error("bad argument");
        return ASD2_ERR_ARG;
    }
    const Geom& G = ctx->geo;
    const size_t n = static_cast<size_t>(G.size());
    MomArgs A = mom_args(fl, 0);
    double *u1, *v1, *u2, *v2;
    ASD2_TRY(ws(ctx, "mom_u1", n, &u1));
    ASD2_TRY(ws(ctx, "mom_v1", n, &v1));
    ASD2_TRY(ws(ctx, "mom_u2", n, &u2));
    ASD2_TRY(ws(ctx, "mom_v2", n, &v2));
    // fresh stage arrays (GetIntermediateVel allocates zero vectors, mac2d.cpp:1146-1147)
    for (double* p : {u1, v1, u2, v2}) ASD2_TRY(zero_async(ctx, p, n * 8));
    auto stage = [&](const double* us, const double* vs, const double* u0, const double* v0, double* uo, double* vo,
                     int mode) -> int {
        A.mode = mode;
        ASD2_TRY(launch_mom(ctx, A, ASD2_GRID_U, false, us, vs, d_ls, d_Hs, u0, nullptr, uo));
        return launch_mom(ctx, A, ASD2_GRID_V, false, us, vs, d_ls, d_Hs, v0, nullptr, vo);
    };
    if (time_order == 0) {
        ASD2_TRY(stage(d_u, d_v, d_u, d_v, d_uhat, d_vhat, 0));
    } else if (time_order == 1) {
        ASD2_TRY(stage(d_u, d_v, d_u, d_v, u1, v1, 0));
        ASD2_TRY(bc_apply_uv(ctx, ASD2_GRID_U, u1));
        ASD2_TRY(bc_apply_uv(ctx, ASD2_GRID_V, v1));
        ASD2_TRY(stage(u1, v1, d_u, d_v, d_uhat, d_vhat, 3));
    } else {
        ASD2_TRY(stage(d_u, d_v, d_u, d_v, u1, v1, 0));
        ASD2_TRY(bc_apply_uv(ctx, ASD2_GRID_U, u1));
        ASD2_TRY(bc_apply_uv(ctx, ASD2_GRID_V, v1));
        ASD2_TRY(stage(u1, v1, d_u, d_v, u2, v2, 1));
        ASD2_TRY(bc_apply_uv(ctx, ASD2_GRID_U, u2));
        ASD2_TRY(bc_apply_uv(ctx, ASD2_GRID_V, v2));
        ASD2_TRY(stage(u2, v2, d_u, d_v, d_uhat, d_vhat, 2));
    }
    return ASD2_OK;
}

int asd2_mac_momentum_term(asd2_ctx* ctx, const asd2_fluid* fl, int grid_kind, int term, const double* d_ls,
                           const double* d_u, const double* d_v, const double* d_Hs, double* d_out) {
    ASD2_TRY(check_ctx(ctx));
    if (!fl || !d_ls || !d_u || !d_v || !d_Hs || !d_out || term < 0 || term > 2 ||
        (grid_kind != ASD2_GRID_U && grid_kind != ASD2_GRID_V)) {
        set_error("bad argument");
        return ASD2_ERR_ARG;
    }
    ASD2_TRY(zero_async(ctx, d_out, static_cast<size_t>(ctx->geo.size()) * 8));
    return launch_mom(ctx, mom_args(fl, 4 + term), grid_kind, false, d_u, d_v, d_ls, d_Hs,
                      grid_kind == ASD2_GRID_U ? d_u : d_v, nullptr, d_out);
}

// the axisymmetric grids need nr, nz >= 3 (the line sweeps solve nr - 1 and nz - 1 unknowns)
static int axi_grid_ok(const asd2_ctx* ctx) {
    if (ctx->geo.nx < 3 || ctx->geo.ny < 3) {
        set_error("axisymmetric momentum: needs nr >= 3 and nz >= 3");
        return ASD2_ERR_ARG;
    }
    return ASD2_OK;
}

int asd2_mac_momentum_term_axisym(asd2_ctx* ctx, const asd2_fluid* fl, int grid_kind, int term, const double* d_ls,
                                  const double* d_u, const double* d_uhat, const double* d_v, const double* d_H,
                                  double* d_out) {
    ASD2_TRY(check_ctx(ctx));
    if (!fl || !d_ls || !d_u || !d_v || !d_H || !d_out || term < 0 || term > 2 ||
        (grid_kind != ASD2_GRID_U && grid_kind != ASD2_GRID_V) || (grid_kind == ASD2_GRID_V && !d_uhat)) {
        set_error("bad argument");
        return ASD2_ERR_ARG;
    }
    ASD2_TRY(axi_grid_ok(ctx));
    if (d_out == d_u || d_out == d_v || d_out == d_uhat) { set_error("d_out aliases an input"); return ASD2_ERR_ARG; }
    ASD2_TRY(zero_async(ctx, d_out, static_cast<size_t>(ctx->geo.size()) * 8));
    const int mode = term == ASD2_MOM_RHS ? 0 : 4 + term;
    // s0 is not read outside the RK modes, but the kernels' clamped loads need a valid address
    ASD2_TRY(launch_mom(ctx, mom_args(fl, mode), grid_kind, true, d_u, d_v, d_ls, d_H,
                        grid_kind == ASD2_GRID_U ? d_u : d_v, d_uhat, d_out));
    if (term == ASD2_MOM_RHS) ASD2_TRY(bc_apply_uv(ctx, grid_kind, d_out));
    return ASD2_OK;
}

int asd2_mac_rhs_u_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_u,
                          const double* d_v, const double* d_H, double* d_rhs) {
    return asd2_mac_momentum_term_axisym(ctx, fl, ASD2_GRID_U, ASD2_MOM_RHS, d_ls, d_u, nullptr, d_v, d_H, d_rhs);
}

int asd2_mac_rhs_v_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_u,
                          const double* d_uhat, const double* d_v, const double* d_H, double* d_rhs) {
    return asd2_mac_momentum_term_axisym(ctx, fl, ASD2_GRID_V, ASD2_MOM_RHS, d_ls, d_u, d_uhat, d_v, d_H, d_rhs);
}

}  // extern "C"
