// levelset.hip -- HJ-WENO5 + SSP-RK3 advection and reinitialisation of the level set on
// gfx950.  Reference: src/levelset2d.cpp.
//
// Kernels keep the reference's FP64 arithmetic.  Branch-deciding quantities (MinMod
// choices, the subcell interface position and its cutoffs, the sign functions) are
// evaluated in the reference's operation order with contraction off, so every branch
// agrees with the reference; the bulk WENO/ENO algebra is reorganised (one division per
// WENO side, reciprocal multiplies) and agrees to rounding.
#include <mutex>

#include "internal.h"
#include "weno.h"

namespace asd2 {

// ------------------------------------------------------------------ WENO5
constexpr int WJ = 64;   // tile width along j (contiguous)
constexpr int WI = 32;   // WENO tile height along i
constexpr int WTH = 256; // 64 x 4 threads
constexpr int WTW = 512; // WENO block: 64 x 8 threads, 4 tile rows each

// Cell-centred velocity of Solve_LevelSet_2D with the P BC on its ghosts (src/levelset2d.cpp:47-57),
// split into its loads and its arithmetic so that a kernel can issue every load of a tile before it
// waits for any.  Along one axis of n interior cells: the source cell of position i (itself inside,
// its mirror on a ghost), and the value from the two face velocities at the source and the next
// face; CUSTOM sides leave the ghost at 0.
__device__ __forceinline__ int centre_src(int i, int g, int n) {
    return i < g ? 2 * g - 1 - i : (i >= n + g ? 2 * n + 2 * g - 1 - i : i);
}
template <int SLO>
__device__ __forceinline__ double centre_vel(const GhostRule& R, int i, int g, int n, double f0, double f1) {
    const double x = (f0 + f1) * 0.5;
    if (i >= g && i < n + g) return x;
    const bool lo = i < g;  // the side's rule, read with constant indices (kernel-argument registers)
    const int kind = lo ? R.kind[SLO] : R.kind[SLO + 1];
    const double tc = lo ? R.two_c[SLO] : R.two_c[SLO + 1];
    if (kind == 2) return 0.0;
    return kind == 0 ? x : (-x + tc);
}

// STAGE 0: write L (HJWENO5_LS_2D); 1..3: SSP-RK3 stages of Solve_LevelSet_2D.
// CENTRED: velocities given cell-centred (with ghosts) instead of on faces.  POW2: dx and dy are
// powers of two.  The WENO derivatives, L and the stage update follow the reference's operation
// order without contraction, so a stage reproduces the reference bit for bit.
// A block owns a WI x WJ tile; the x fluxes phi u_c need 3 halo rows above and below it, the y fluxes
// phi v_c 3 halo columns on each side.  Every global load of a thread (its NR tile cells: phi, the
// velocity sources, phi^n; one x-halo and one y-halo cell) is issued at addresses clamped into the
// array before the first value is used.
template <int STAGE, bool CENTRED, bool POW2>
__global__ __launch_bounds__(WTW) void k_weno(Geom G, GhostRule R, const double* __restrict__ phi,
                                              const double* phin, const double* __restrict__ u,
                                              const double* __restrict__ v, double* out, double dt,
                                              int ghosts) {
    __shared__ double sFx[WI + 6][WJ];
    __shared__ double sFy[WI][WJ + 6];
    const int tj = threadIdx.x & (WJ - 1);
    const int ti = threadIdx.x / WJ;
    const int g = G.g;
    const int j0 = g + blockIdx.x * WJ;
    const int i0 = g + blockIdx.y * WI;
    const int iend = G.nx + g, jend = G.ny + g;
    const int ipad = G.nx + 2 * g, jpad = G.ny + 2 * g;

    // Each thread owns the output rows o = ti + 8k of its column; phi and the cell-centred velocities
    // stay in registers for the stage update after the flux phase.
    constexpr int NR = WI / (WTW / WJ);
    double phiR[NR], ua[NR], ub[NR], va[NR], vb[NR], pn[NR];
    int okR = 0;
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int i = i0 + ti + (WTW / WJ) * k, j = j0 + tj;
        const bool ok = i < ipad && j < jpad;
        okR |= ok ? (1 << k) : 0;
        const int ic = ok ? i : g, jc = ok ? j : g;
        phiR[k] = phi[G.idx(ic, jc)];
        if (CENTRED) {
            ua[k] = u[G.idx(ic, jc)];
            va[k] = v[G.idx(ic, jc)];
            ub[k] = vb[k] = 0.0;
        } else {
            const int mi = centre_src(ic, g, G.nx), mj = centre_src(jc, g, G.ny);
            ua[k] = u[G.idx(mi, jc)];
            ub[k] = u[G.idx(mi + 1, jc)];
            va[k] = v[G.idx(ic, mj)];
            vb[k] = v[G.idx(ic, mj + 1)];
        }
        pn[k] = (STAGE >= 2) ? phin[G.idx(ic, jc)] : 0.0;
    }
    // x-flux halo: rows i0 - 3 .. i0 - 1 and i0 + WI .. i0 + WI + 2 (one cell per thread < 6 WJ)
    const int ex = threadIdx.x;
    const int hxr = (ex / WJ) < 3 ? (ex / WJ) : WI + (ex / WJ);
    const int hxi = i0 - 3 + hxr, hxj = j0 + (ex & (WJ - 1));
    const bool hxok = ex < 6 * WJ && hxj < jend && hxi < ipad;
    const int hxic = hxok ? hxi : g, hxjc = hxok ? hxj : g;
    const double hxp = phi[G.idx(hxic, hxjc)];
    const int hxm = CENTRED ? hxic : centre_src(hxic, g, G.nx);
    const double hx0 = u[G.idx(hxm, hxjc)];
    const double hx1 = CENTRED ? 0.0 : u[G.idx(hxm + 1, hxjc)];
    // y-flux halo: columns j0 - 3 .. j0 - 1 and j0 + WJ .. j0 + WJ + 2 (one cell per thread < 6 WI)
    const int ey = threadIdx.x;
    const int hyr = ey / 6, hyh = ey - hyr * 6;
    const int hyi = i0 + hyr, hyj = j0 - 3 + (hyh < 3 ? hyh : WJ + hyh);
    const bool hyok = ey < 6 * WI && hyi < iend && hyj < jpad;
    const int hyic = hyok ? hyi : g, hyjc = hyok ? hyj : g;
    const double hyp = phi[G.idx(hyic, hyjc)];
    const int hym = CENTRED ? hyjc : centre_src(hyjc, g, G.ny);
    const double hy0 = v[G.idx(hyic, hym)];
    const double hy1 = CENTRED ? 0.0 : v[G.idx(hyic, hym + 1)];

    double ucR[NR], vcR[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int o = ti + (WTW / WJ) * k, i = i0 + o, j = j0 + tj;
        const bool ok = (okR >> k) & 1;
        const double ph = ok ? phiR[k] : 0.0;
        const double uc = !ok ? 0.0 : CENTRED ? ua[k] : centre_vel<0>(R, i, g, G.nx, ua[k], ub[k]);
        const double vc = !ok ? 0.0 : CENTRED ? va[k] : centre_vel<2>(R, j, g, G.ny, va[k], vb[k]);
        sFx[o + 3][tj] = (j < jend && i < ipad) ? ph * uc : 0.0;
        sFy[o][tj + 3] = (i < iend && j < jpad) ? ph * vc : 0.0;
        phiR[k] = ph;
        ucR[k] = uc;
        vcR[k] = vc;
    }
    if (ex < 6 * WJ) {
        const double vel = CENTRED ? hx0 : centre_vel<0>(R, hxi, g, G.nx, hx0, hx1);
        sFx[hxr][ex & (WJ - 1)] = hxok ? hxp * vel : 0.0;
    }
    if (ey < 6 * WI) {
        const double vel = CENTRED ? hy0 : centre_vel<2>(R, hyj, g, G.ny, hy0, hy1);
        sFy[hyr][hyh < 3 ? hyh : WJ + hyh] = hyok ? hyp * vel : 0.0;
    }
    __syncthreads();

    // one row at a time (interleaving the rows' division chains costs more registers than it saves)
    const double idx_ = 1.0 / G.dx, idy_ = 1.0 / G.dy;
    double val[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int r = ti + (WTW / WJ) * k;
        if (i0 + r >= iend || j0 + tj >= jend) continue;
        double lxp, lxm, lyp, lym;
        {
            const double f0 = sFx[r][tj], f1 = sFx[r + 1][tj], f2 = sFx[r + 2][tj], f3 = sFx[r + 3][tj];
            const double f4 = sFx[r + 4][tj], f5 = sFx[r + 5][tj], f6 = sFx[r + 6][tj];
            weno5_ref<POW2>(f1 - f0, f2 - f1, f3 - f2, f4 - f3, f5 - f4, f6 - f5, G.dx, idx_, lxp, lxm);
        }
        {
            const double f0 = sFy[r][tj], f1 = sFy[r][tj + 1], f2 = sFy[r][tj + 2], f3 = sFy[r][tj + 3];
            const double f4 = sFy[r][tj + 4], f5 = sFy[r][tj + 5], f6 = sFy[r][tj + 6];
            weno5_ref<POW2>(f1 - f0, f2 - f1, f3 - f2, f4 - f3, f5 - f4, f6 - f5, G.dy, idy_, lyp, lym);
        }
        const double ax = ucR[k], ay = vcR[k], phc = phiR[k];
        {
#pragma clang fp contract(off)
            const double L =
                0.5 * (lxp + lxm) + 0.5 * (lyp + lym) - ax * (0.5 * (lxp - lxm)) - ay * (0.5 * (lyp - lym));
            if (STAGE == 0) val[k] = L;
            else if (STAGE == 1) val[k] = phc - dt * L;
            else if (STAGE == 2) val[k] = 0.75 * pn[k] + 0.25 * phc - 0.25 * dt * L;
            else val[k] = 1.0 / 3.0 * pn[k] + 2.0 / 3.0 * phc - 2.0 / 3.0 * dt * L;
        }
    }
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const int i = i0 + ti + (WTW / WJ) * k, j = j0 + tj;
        if (i >= iend || j >= jend) continue;
        out[G.idx(i, j)] = val[k];
        if (ghosts) write_p_ghosts(G, R, out, i, j, val[k]);
    }
}

// interior cell-centred velocities of Solve_LevelSet_2D (src/levelset2d.cpp:50-54); used
// when a side's ghost rule cannot be evaluated inside the stage kernel (LS-free sides)
__global__ __launch_bounds__(256) void k_centre_velocity(Geom G, const double* __restrict__ u,
                                                         const double* __restrict__ v, double* __restrict__ uc,
                                                         double* __restrict__ vc) {
    const int j = G.g + blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = G.g + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + G.g || j >= G.ny + G.g) return;
    const int64_t c = G.idx(i, j);
    uc[c] = (u[c] + u[G.idx(i + 1, j)]) * 0.5;
    vc[c] = (v[c] + v[G.idx(i, j + 1)]) * 0.5;
}

// ------------------------------------------------------------------ ENO / subcell
constexpr int RJ = 64, RI = 16, RH = 2;  // reinit tile and halo
constexpr int RW = RJ + 2 * RH;

// Loads the TH x TW tiles of a and b whose top-left cell is (ib, jb) into sa and sb, 0 outside the
// padded array.  Every load of a thread is issued before the first LDS write (clamped addresses,
// no branch around a load), so a block waits for one memory latency instead of one per pass of a
// guarded load loop.
template <int TH, int TW, int NT>
__device__ __forceinline__ void tile_load2(const Geom& G, const double* __restrict__ a, const double* __restrict__ b,
                                           double (*sa)[TW], double (*sb)[TW], int ib, int jb) {
    constexpr int NE = TH * TW, NL = (NE + NT - 1) / NT;
    const int ipad = G.nx + 2 * G.g, jpad = G.ny + 2 * G.g;
    double va[NL], vb[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const int e = threadIdx.x + k * NT;
        const int r = e / TW, c = e - r * TW;
        const int i = ib + r, j = jb + c;
        const bool in = e < NE && i >= 0 && i < ipad && j >= 0 && j < jpad;
        const int64_t o = in ? G.idx(i, j) : 0;
        const double x = a[o], y = b[o];
        va[k] = in ? x : 0.0;
        vb[k] = in ? y : 0.0;
    }
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const int e = threadIdx.x + k * NT;
        if (NE % NT == 0 || e < NE) {
            const int r = e / TW, c = e - r * TW;
            sa[r][c] = va[k];
            sb[r][c] = vb[k];
        }
    }
}

// 2nd-order ENO one-sided differences with MinMod (src/levelset2d.cpp:573-593):
// the MinMod picks on undivided second differences (same choice as on divided ones)
// and the result is scaled once.  x0..x4 = phi[k-2..k+2] along one axis.
__device__ __forceinline__ void eno2(double xm2, double xm1, double x0, double xp1, double xp2,
                                     double inv_h, double& dp, double& dm) {
    const double dd0 = xm1 - 2.0 * x0 + xp1;
    const double ddp = x0 - 2.0 * xp1 + xp2;
    const double ddm = xm2 - 2.0 * xm1 + x0;
    dp = inv_h * ((xp1 - x0) - 0.5 * minmod(dd0, ddp));
    dm = inv_h * ((x0 - xm1) + 0.5 * minmod(dd0, ddm));
}

// Min-Gibou subcell interface distance and one-sided derivative
// (src/levelset2d.cpp:655-707), reference operation order, no contraction.
// I* = phi0 along the axis (k-2..k+2), l0 = current phi at k; plus = derivative towards k+1.
__device__ __noinline__ double subcell_plus(double Im1, double I0, double Ip1, double Ip2, double l0, double h) {
#pragma clang fp contract(off)
    const double DD = minmod((Im1 - 2.0 * I0 + Ip1) / (h * h), (I0 - 2.0 * Ip1 + Ip2) / (h * h));
    const double c2 = 0.5 * DD;
    const double c1 = (Ip1 - I0) / h;
    const double c0 = (Ip1 + I0) * 0.5 - c2 * h * h * 0.25;
    double nd;
    if (fabs(c2) > 1.0e-10) nd = 0.5 * h + (-c1 - isign(I0) * sqrt(c1 * c1 - 4.0 * c2 * c0)) / (2.0 * c2);
    else nd = 0.5 * h - c0 / c1;
    double d = -l0 / nd - 0.5 * nd * DD;
    if (fabs(nd) < h * 0.01) d = -0.5 * nd * DD;
    return d;
}

__device__ __noinline__ double subcell_minus(double Im2, double Im1, double I0, double Ip1, double l0, double h) {
#pragma clang fp contract(off)
    const double DD = minmod((Im1 - 2.0 * I0 + Ip1) / (h * h), (Im2 - 2.0 * Im1 + I0) / (h * h));
    const double c2 = 0.5 * DD;
    const double c1 = (I0 - Im1) / h;
    const double c0 = (I0 + Im1) * 0.5 - c2 * h * h * 0.25;
    double nd;
    if (fabs(c2) > 1.0e-10) nd = 0.5 * h - (-c1 + isign(I0) * sqrt(c1 * c1 - 4.0 * c2 * c0)) / (2.0 * c2);
    else nd = 0.5 * h + c0 / c1;
    double d = l0 / nd + 0.5 * nd * DD;
    if (fabs(nd) < h * 0.01) d = 0.5 * nd * DD;
    return d;
}

// Godunov |grad phi| (src/levelset2d.cpp:768-782)
__device__ __forceinline__ double godunov(double s0, double dPX, double dMX, double dPY, double dMY) {
    double a, b;
    if (s0 >= 0) {
        const double px = smin(dPX, 0.0), mx = smax(dMX, 0.0), py = smin(dPY, 0.0), my = smax(dMY, 0.0);
        a = smax(px * px, mx * mx);
        b = smax(py * py, my * my);
    } else {
        const double px = smax(dPX, 0.0), mx = smin(dMX, 0.0), py = smax(dPY, 0.0), my = smin(dMY, 0.0);
        a = smax(px * px, mx * mx);
        b = smax(py * py, my * my);
    }
    return sqrt(a + b);
}

// |grad phi| at tile position (r, c) (offset by the halo) from tiles of phi and phi0.
template <bool SUBCELL>
__device__ __forceinline__ double absgrad_at(const double (*sL)[RW], const double (*sI)[RW], int r, int c,
                                             double dx, double dy, double idx_, double idy_) {
    const int R = r + RH, C = c + RH;
    double dPX, dMX, dPY, dMY;
    const double l0 = sL[R][C];
    eno2(sL[R - 2][C], sL[R - 1][C], l0, sL[R + 1][C], sL[R + 2][C], idx_, dPX, dMX);
    eno2(sL[R][C - 2], sL[R][C - 1], l0, sL[R][C + 1], sL[R][C + 2], idy_, dPY, dMY);
    const double I0 = sI[R][C];
    if (SUBCELL) {
        const double Ixm2 = sI[R - 2][C], Ixm1 = sI[R - 1][C], Ixp1 = sI[R + 1][C], Ixp2 = sI[R + 2][C];
        const double Iym2 = sI[R][C - 2], Iym1 = sI[R][C - 1], Iyp1 = sI[R][C + 1], Iyp2 = sI[R][C + 2];
        if (I0 * Ixp1 < 0) dPX = subcell_plus(Ixm1, I0, Ixp1, Ixp2, l0, dx);
        if (I0 * Ixm1 < 0) dMX = subcell_minus(Ixm2, Ixm1, I0, Ixp1, l0, dx);
        if (I0 * Iyp1 < 0) dPY = subcell_plus(Iym1, I0, Iyp1, Iyp2, l0, dy);
        if (I0 * Iym1 < 0) dMY = subcell_minus(Iym2, Iym1, I0, Iyp1, l0, dy);
    }
    return godunov(I0, dPX, dMX, dPY, dMY);
}

// smoothed sign of phi0 (src/levelset2d.cpp:914-923), reference operation order
__device__ __forceinline__ double smoothed_sign(double x, double eps) {
#pragma clang fp contract(off)
    return x / sqrt(x * x + eps * eps);
}

// mode 0: out = |grad phi| (subcell) ; 1: out = |grad phi| (ENO)
template <bool SUBCELL>
__global__ __launch_bounds__(WTH) void k_absgrad(Geom G, const double* __restrict__ ls,
                                                 const double* __restrict__ ls0, double* __restrict__ out) {
    __shared__ double sL[RI + 2 * RH][RW];
    __shared__ double sI[RI + 2 * RH][RW];
    const int j0 = G.g + blockIdx.x * RJ, i0 = G.g + blockIdx.y * RI;
    tile_load2<RI + 2 * RH, RW, WTH>(G, ls, ls0, sL, sI, i0 - RH, j0 - RH);
    __syncthreads();
    const int tj = threadIdx.x & (RJ - 1), ti = threadIdx.x / RJ;
    const double idx_ = 1.0 / G.dx, idy_ = 1.0 / G.dy;
    for (int r = ti; r < RI; r += WTH / RJ) {
        const int i = i0 + r, j = j0 + tj;
        if (i >= G.nx + G.g || j >= G.ny + G.g) continue;
        out[G.idx(i, j)] = absgrad_at<SUBCELL>(sL, sI, r, tj, G.dx, G.dy, idx_, idy_);
    }
}

// MinRK2 substage A: LS1 = phi - dtau S(phi0) (|grad phi|_sub - 1)   (levelset2d.cpp:453-460)
// MinRK2 substage B: LS2 = LS1 - dtau S (|grad LS1|_sub - 1); phi = (phi + LS2)/2  (:462-481)
// Sussman (SUBCELL = false, sign(phi0)): A writes LS1 without ghosts (:344-348),
// B writes tmp = (phi + LS2)/2 with ghosts (:351-363).
// SUB 2 (first-time Sussman, :405-421): LS2 = phi - dtau sign(phi0) (|grad LS1| - 1),
// phi = (phi + LS2)/2, i.e. the second stage restarts from phi instead of LS1.  SUB 3: that LS2
// alone (the average left to k_ls_average, for grids whose P BC must run on LS2 first).
template <bool SUBCELL, int SUB>
__global__ __launch_bounds__(WTH) void k_reinit(Geom G, GhostRule R, const double* __restrict__ src,
                                                const double* __restrict__ ls0, const double* phi_c,
                                                double* out, int ghosts) {
    __shared__ double sL[RI + 2 * RH][RW];
    __shared__ double sI[RI + 2 * RH][RW];
    const int j0 = G.g + blockIdx.x * RJ, i0 = G.g + blockIdx.y * RI;
    tile_load2<RI + 2 * RH, RW, WTH>(G, src, ls0, sL, sI, i0 - RH, j0 - RH);
    __syncthreads();
    const int tj = threadIdx.x & (RJ - 1), ti = threadIdx.x / RJ;
    const double idx_ = 1.0 / G.dx, idy_ = 1.0 / G.dy;
    const double eps = smin(G.dx, G.dy), adt = eps * 0.5;
    for (int r = ti; r < RI; r += WTH / RJ) {
        const int i = i0 + r, j = j0 + tj;
        if (i >= G.nx + G.g || j >= G.ny + G.g) continue;
        const double a = absgrad_at<SUBCELL>(sL, sI, r, tj, G.dx, G.dy, idx_, idy_);
        const double I0 = sI[r + RH][tj + RH];
        double val;
        {
#pragma clang fp contract(off)
            const double S = SUBCELL ? smoothed_sign(I0, eps) : static_cast<double>(isign(I0));
            const double base = (SUB == 2 || SUB == 3) ? phi_c[G.idx(i, j)] : sL[r + RH][tj + RH];
            const double x = base - adt * S * (a - 1.0);
            if (SUB == 0 || SUB == 3) val = x;
            else val = 0.5 * (phi_c[G.idx(i, j)] + x);
        }
        out[G.idx(i, j)] = val;
        if (ghosts) write_p_ghosts(G, R, out, i, j, val);
    }
}

// phi = (phi + LS2) / 2 on the interior cells (the reinitialisations' average, taken after the P BC
// of LS2 where a periodic east side makes that BC change interior rows)
__global__ __launch_bounds__(256) void k_ls_average(Geom G, double* __restrict__ phi, const double* __restrict__ ls2) {
    const int j = G.g + blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = G.g + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + G.g || j >= G.ny + G.g) return;
    const int64_t c = G.idx(i, j);
    phi[c] = 0.5 * (phi[c] + ls2[c]);
}

// |grad phi| with the subcell fix from a phi tile and a phi0 tile of different halos: (RL, CL) and
// (RI_, CI) are the cell's positions in the two tiles.  The same arithmetic as absgrad_at<true>.
template <int LW, int IW>
__device__ __forceinline__ double absgrad_sub2(const double (*sL)[LW], int RL, int CL, const double (*sI)[IW], int RI_,
                                               int CI, double dx, double dy, double idx_, double idy_) {
    double dPX, dMX, dPY, dMY;
    const double l0 = sL[RL][CL];
    eno2(sL[RL - 2][CL], sL[RL - 1][CL], l0, sL[RL + 1][CL], sL[RL + 2][CL], idx_, dPX, dMX);
    eno2(sL[RL][CL - 2], sL[RL][CL - 1], l0, sL[RL][CL + 1], sL[RL][CL + 2], idy_, dPY, dMY);
    const double I0 = sI[RI_][CI];
    const double Ixm2 = sI[RI_ - 2][CI], Ixm1 = sI[RI_ - 1][CI], Ixp1 = sI[RI_ + 1][CI], Ixp2 = sI[RI_ + 2][CI];
    const double Iym2 = sI[RI_][CI - 2], Iym1 = sI[RI_][CI - 1], Iyp1 = sI[RI_][CI + 1], Iyp2 = sI[RI_][CI + 2];
    if (I0 * Ixp1 < 0) dPX = subcell_plus(Ixm1, I0, Ixp1, Ixp2, l0, dx);
    if (I0 * Ixm1 < 0) dMX = subcell_minus(Ixm2, Ixm1, I0, Ixp1, l0, dx);
    if (I0 * Iyp1 < 0) dPY = subcell_plus(Iym1, I0, Iyp1, Iyp2, l0, dy);
    if (I0 * Iym1 < 0) dMY = subcell_minus(Iym2, Iym1, I0, Iyp1, l0, dy);
    return godunov(I0, dPX, dMX, dPY, dMY);
}

// One whole MinRK2 iteration (src/levelset2d.cpp:451-481) per launch, for grids whose P ghost
// rules the kernels can evaluate (no LS-free side).  A block owns an FI x FJ tile: it loads phi
// and phi0 with a 4-cell halo, runs substage A (LS1 = phi - dtau S (|grad phi|_sub - 1)) on the
// tile and on the 2-cell cross-shaped ring the ENO/subcell stencils of substage B read, fills
// the ring's ghost cells from LS1's interior values by the P ghost rule (what ApplyBC_P_2D does to
// LS1 between the substages), then runs substage B and phi' = (phi + LS2) / 2 on the tile and
// writes phi' with its ghosts.  phi' must not alias phi: neighbouring blocks read phi's halo.
// (CUSTOM sides keep the caller's ghosts of phi, which a ping-ponged phi' would not carry: those
// grids take the two-launch path.)
// Per cell and iteration: phi and phi0 read, phi' written (24 B), against 56 B for the two
// substage launches; the ring costs (FI + 4) FJ + 4 FI substage-A evaluations per FI FJ cells.
constexpr int FJ = 64, FH = 4;
constexpr int RFI = 32, RFT = 512;  // tile height and threads of the launches

template <int FI, int NT>
__global__ __launch_bounds__(NT) void k_reinit_minrk2(Geom G, GhostRule R, const double* __restrict__ phi,
                                                      const double* __restrict__ ls0, double* __restrict__ out) {
    constexpr int PW = FJ + 2 * FH, PH = FI + 2 * FH;  // phi / phi0 tiles: halo 4
    constexpr int LW = FJ + 4, LH = FI + 4;            // LS1 tile: halo 2
    __shared__ double sP[PH][PW];
    __shared__ double sI[PH][PW];
    __shared__ double sL[LH][LW];
    const int g = G.g;
    const int j0 = g + blockIdx.x * FJ, i0 = g + blockIdx.y * FI;
    const int iend = min(i0 + FI, G.nx + g), jend = min(j0 + FJ, G.ny + g);
    tile_load2<PH, PW, NT>(G, phi, ls0, sP, sI, i0 - FH, j0 - FH);
    __syncthreads();
    const double idx_ = 1.0 / G.dx, idy_ = 1.0 / G.dy;
    const double eps = smin(G.dx, G.dy), adt = eps * 0.5;
    // substage A on the interior cells of the cross-shaped ring + tile
    for (int e = threadIdx.x; e < LH * LW; e += NT) {
        const int r = e / LW, c = e - r * LW;
        const int i = i0 - 2 + r, j = j0 - 2 + c;
        const bool need = (i >= i0 - 2 && i < iend + 2 && j >= j0 && j < jend) ||
                          (i >= i0 && i < iend && j >= j0 - 2 && j < jend + 2);
        if (!need || i < g || i >= G.nx + g || j < g || j >= G.ny + g) continue;
        const double a = absgrad_sub2<PW, PW>(sP, r + 2, c + 2, sI, r + 2, c + 2, G.dx, G.dy, idx_, idy_);
        const double I0 = sI[r + 2][c + 2];
        double x;
        {
#pragma clang fp contract(off)
            const double S = smoothed_sign(I0, eps);
            x = sP[r + 2][c + 2] - adt * S * (a - 1.0);
        }
        sL[r][c] = x;
    }
    __syncthreads();
    // LS1's ghost cells on the ring (one coordinate outside the interior; the mirror is interior
    // and was computed above)
    for (int e = threadIdx.x; e < LH * LW; e += NT) {
        const int r = e / LW, c = e - r * LW;
        const int i = i0 - 2 + r, j = j0 - 2 + c;
        const bool need = (i >= i0 - 2 && i < iend + 2 && j >= j0 && j < jend) ||
                          (i >= i0 && i < iend && j >= j0 - 2 && j < jend + 2);
        if (!need) continue;
        int side = -1, mi = i, mj = j;
        if (i < g) { side = 0; mi = 2 * g - 1 - i; }
        else if (i >= G.nx + g) { side = 1; mi = 2 * G.nx + 2 * g - 1 - i; }
        else if (j < g) { side = 2; mj = 2 * g - 1 - j; }
        else if (j >= G.ny + g) { side = 3; mj = 2 * G.ny + 2 * g - 1 - j; }
        if (side < 0) continue;
        const double m = sL[mi - (i0 - 2)][mj - (j0 - 2)];
        // the side's rule with constant indices (kernel-argument registers, no memory load)
        const int kind = side == 0 ? R.kind[0] : side == 1 ? R.kind[1] : side == 2 ? R.kind[2] : R.kind[3];
        const double tc = side == 0 ? R.two_c[0] : side == 1 ? R.two_c[1] : side == 2 ? R.two_c[2] : R.two_c[3];
        sL[r][c] = kind == 0 ? m : kind == 1 ? (-m + tc) : 0.0;
    }
    __syncthreads();
    // substage B and the average on the tile
    for (int e = threadIdx.x; e < FI * FJ; e += NT) {
        const int r = e / FJ, c = e - r * FJ;
        const int i = i0 + r, j = j0 + c;
        if (i >= iend || j >= jend) continue;
        const double a = absgrad_sub2<LW, PW>(sL, r + 2, c + 2, sI, r + FH, c + FH, G.dx, G.dy, idx_, idy_);
        const double I0 = sI[r + FH][c + FH];
        double val;
        {
#pragma clang fp contract(off)
            const double S = smoothed_sign(I0, eps);
            const double x = sL[r + 2][c + 2] - adt * S * (a - 1.0);
            val = 0.5 * (sP[r + FH][c + FH] + x);
        }
        out[G.idx(i, j)] = val;
        write_p_ghosts(G, R, out, i, j, val);
    }
}

// Sussman's volume constraint (src/levelset2d.cpp:27-41 mask, :787-893 correction),
// reference operation order; phi += correction, ghosts rewritten.
constexpr int SJ = 64, SI = 16, SHL = 1, SW = SJ + 2 * SHL;

__device__ __noinline__ double sussman_correction(const double (*T)[SW], const double (*I)[SW], int R, int C,
                                                  double dx, double dy) {
#pragma clang fp contract(off)
    const double eps = smin(dx, dy), thick = 1.5 * eps, adt = eps * 0.5;
    double F[3][3], IF[3][3], dA[3][3], Hp[3][3], Lm[3][3], f[3][3];
#define L_(a, b) T[R + (a)][C + (b)]
#define I_(a, b) I[R + (a)][C + (b)]
    F[0][0] = 0.25 * (L_(0, 0) + L_(-1, 0) + L_(-1, -1) + L_(0, -1));
    F[0][1] = 0.5 * (L_(0, -1) + L_(0, 0));
    F[0][2] = 0.25 * (L_(0, 0) + L_(-1, 0) + L_(-1, 1) + L_(0, 1));
    F[1][0] = 0.5 * (L_(0, -1) + L_(0, 0));
    F[1][1] = L_(0, 0);
    F[1][2] = 0.5 * (L_(0, 0) + L_(0, 1));
    F[2][0] = 0.25 * (L_(0, 0) + L_(1, 0) + L_(1, -1) + L_(0, -1));
    F[2][1] = 0.5 * (L_(0, 0) + L_(1, 0));
    F[2][2] = 0.25 * (L_(0, 0) + L_(1, 0) + L_(1, 1) + L_(0, 1));
    IF[0][0] = 0.25 * (I_(0, 0) + I_(-1, 0) + I_(-1, -1) + I_(0, -1));
    IF[0][1] = 0.5 * (I_(0, -1) + I_(0, 0));
    IF[0][2] = 0.25 * (I_(0, 0) + I_(-1, 0) + I_(-1, 1) + I_(0, 1));
    IF[1][0] = 0.5 * (I_(0, -1) + I_(0, 0));
    IF[1][1] = I_(0, 0);
    IF[1][2] = 0.5 * (I_(0, 0) + I_(0, 1));
    IF[2][0] = 0.25 * (I_(0, 0) + I_(1, 0) + I_(1, -1) + I_(0, -1));
    IF[2][1] = 0.5 * (I_(0, 0) + I_(1, 0));
    IF[2][2] = 0.25 * (I_(0, 0) + I_(1, 0) + I_(1, 1) + I_(0, 1));
    auto sq = [](double x) { return x * x; };
    dA[0][0] = sqrt(sq(((I_(0, 0) + I_(0, -1)) * 0.5 - 0.5 * (I_(-1, 0) + I_(-1, -1))) / dx)
                  + sq((IF[0][1] - 0.5 * (I_(0, -1) + I_(-1, -1))) / dy));
    dA[0][1] = sqrt(sq((IF[1][1] - I_(-1, 0)) / dx) + sq((IF[2][0] - IF[0][0]) / dy));
    dA[0][2] = sqrt(sq((IF[1][2] - -0.5 * (I_(-1, 0) + I_(-1, 1))) / dx)
                  + sq((0.5 * (I_(-1, 1) + I_(0, 1)) - IF[0][1]) / dy));
    dA[1][0] = sqrt(sq((IF[2][0] - IF[0][0]) / dx) + sq((IF[1][1] - 0.5 * (I_(0, -1) + I_(0, 0))) / dy));
    dA[1][1] = sqrt(sq((IF[2][1] - IF[0][1]) / dx) + sq((IF[1][2] - IF[1][0]) / dy));
    dA[1][2] = sqrt(sq((IF[2][2] - IF[0][2]) / dx) + sq((I_(0, 1) - IF[1][1]) / dy));
    dA[2][0] = sqrt(sq((0.5 * (I_(1, 0) + I_(1, -1)) - IF[1][0]) / dx)
                  + sq((IF[2][1] - 0.5 * (I_(0, -1) + I_(1, -1))) / dy));
    dA[2][1] = sqrt(sq((I_(1, 0) - IF[1][1]) / (2.0 * dx)) + sq((IF[2][2] - IF[2][0]) / dy));
    dA[2][2] = sqrt(sq((0.5 * (I_(1, 0) + I_(1, 1)) - IF[1][1]) / dx)
                  + sq((0.5 * (I_(0, 1) + I_(1, 1)) - IF[2][1]) / dy));
#undef L_
#undef I_
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            if (fabs(F[a][b]) > thick) Hp[a][b] = 0.0;
            else Hp[a][b] = 0.5 / eps + 0.5 / eps * cos(M_PI * F[a][b] / eps);
            Lm[a][b] = (F[a][b] - IF[a][b]) / adt;
            f[a][b] = Hp[a][b] * dA[a][b];
        }
    double s1 = 0.0, s2 = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            s1 += Hp[a][b] * Lm[a][b];
            s2 += Hp[a][b] * f[a][b];
        }
    s1 += 15.0 * Hp[1][1] * Lm[1][1];
    s2 += 15.0 * Hp[1][1] * f[1][1];
    return 0.0 - adt * f[1][1] * s1 / s2;
}

// add = 1: ls += correction (in place) with ghosts rewritten (one Sussman iteration);
// add = 0: out = correction, with the mask taken from hd (GetSussmanReinitConstraint).
__global__ __launch_bounds__(WTH) void k_sussman_constraint(Geom G, GhostRule R, const double* __restrict__ tmp,
                                                            const double* __restrict__ ls0, const double* hd,
                                                            double* out, int add, int ghosts) {
    __shared__ double sT[SI + 2 * SHL][SW];
    __shared__ double sI[SI + 2 * SHL][SW];
    const int j0 = G.g + blockIdx.x * SJ, i0 = G.g + blockIdx.y * SI;
    const int ipad = G.nx + 2 * G.g, jpad = G.ny + 2 * G.g;
    for (int e = threadIdx.x; e < (SI + 2 * SHL) * SW; e += WTH) {
        const int r = e / SW, c = e - r * SW;
        const int i = i0 - SHL + r, j = j0 - SHL + c;
        const bool in = (i < ipad && j < jpad);
        sT[r][c] = in ? tmp[G.idx(i, j)] : 0.0;
        sI[r][c] = in ? ls0[G.idx(i, j)] : 0.0;
    }
    __syncthreads();
    const int tj = threadIdx.x & (SJ - 1), ti = threadIdx.x / SJ;
    const double eps = smin(G.dx, G.dy);
    for (int r = ti; r < SI; r += WTH / SJ) {
        const int i = i0 + r, j = j0 + tj;
        if (i >= G.nx + G.g || j >= G.ny + G.g) continue;
        const int64_t c = G.idx(i, j);
        const double t = sT[r + SHL][tj + SHL];
        double mask;
        if (hd) mask = hd[c];
        else if (fabs(t) > eps) mask = 0.0;
        else {
#pragma clang fp contract(off)
            mask = 0.5 * (1.0 / eps + cos(M_PI * t / eps));
        }
        const double cons = (mask != 0.0) ? sussman_correction(sT, sI, r + SHL, tj + SHL, G.dx, G.dy) : 0.0;
        if (add) {
            const double val = out[c] + cons;
            out[c] = val;
            if (ghosts) write_p_ghosts(G, R, out, i, j, val);
        } else {
            out[c] = cons;
        }
    }
}

// UpdateHeavisideFuncDeriv (what = 0) and GetSmoothedSignFunc (what = 1) over the padded array
__global__ __launch_bounds__(256) void k_pointwise(int64_t n, int what, double eps, const double* __restrict__ in,
                                                   double* __restrict__ out) {
#pragma clang fp contract(off)
    for (int64_t k = blockIdx.x * 256ll + threadIdx.x; k < n; k += static_cast<int64_t>(gridDim.x) * 256) {
        const double x = in[k];
        double y;
        if (what == 0) y = (fabs(x) > eps) ? 0.0 : 0.5 * (1.0 / eps + cos(M_PI * x / eps));
        else y = smoothed_sign(x, eps);
        out[k] = y;
    }
}

// one WENO5 line: FP, FM on [g, n + g) from F[0, n + 2g)
__global__ __launch_bounds__(256) void k_unit_weno5(const double* __restrict__ F, double* __restrict__ FP,
                                                    double* __restrict__ FM, double d, int n, int g) {
    const int i = g + blockIdx.x * 256 + threadIdx.x;
    if (i >= n + g) return;
    double f[7];
    for (int k = 0; k < 7; k++) f[k] = F[i - 3 + k];
    double fp, fm;
    weno5_ref<false>(f[1] - f[0], f[2] - f[1], f[3] - f[2], f[4] - f[3], f[5] - f[4], f[6] - f[5], d, 1.0 / d, fp, fm);
    FP[i] = fp;
    FM[i] = fm;
}

// HJWENO5_DerivAbsLS_2D (src/levelset2d.cpp:488-557): WENO5 one-sided derivatives of phi
// along both axes, Godunov |grad phi| with the branch chosen by phi0 >= 0.  Off the hot path
// (the reference's reinitialisations use ENO2); one thread per interior cell.
__global__ __launch_bounds__(256) void k_weno_absgrad(Geom G, const double* __restrict__ ls,
                                                      const double* __restrict__ ls0, double* __restrict__ out) {
    const int j = G.g + blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = G.g + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + G.g || j >= G.ny + G.g) return;
    double f[7], dPX, dMX, dPY, dMY;
    for (int k = 0; k < 7; k++) f[k] = ls[G.idx(i - 3 + k, j)];
    weno5_ref<false>(f[1] - f[0], f[2] - f[1], f[3] - f[2], f[4] - f[3], f[5] - f[4], f[6] - f[5], G.dx, 1.0 / G.dx,
                     dPX, dMX);
    for (int k = 0; k < 7; k++) f[k] = ls[G.idx(i, j - 3 + k)];
    weno5_ref<false>(f[1] - f[0], f[2] - f[1], f[3] - f[2], f[4] - f[3], f[5] - f[4], f[6] - f[5], G.dy, 1.0 / G.dy,
                     dPY, dMY);
    out[G.idx(i, j)] = godunov(ls0[G.idx(i, j)], dPX, dMX, dPY, dMY);
}

struct Launch {
    dim3 grid, block;
};

static Launch tiles(const Geom& G, int tj, int ti, int threads = WTH) {
    return Launch{dim3(cdiv(G.ny, tj), cdiv(G.nx, ti)), dim3(threads)};
}

// one WENO stage launch; the power-of-two variant when both spacings are powers of two
template <int STAGE, bool CENTRED>
void launch_weno(asd2_ctx* ctx, const GhostRule& R, const double* phi, const double* phin, const double* u,
                 const double* v, double* out, double dt, int ghosts) {
    const Geom& G = ctx->geo;
    const Launch L = tiles(G, WJ, WI, WTW);
    if (is_pow2(G.dx) && is_pow2(G.dy))
        hipLaunchKernelGGL((k_weno<STAGE, CENTRED, true>), L.grid, L.block, 0, ctx->stream, G, R, phi, phin, u, v, out,
                           dt, ghosts);
    else
        hipLaunchKernelGGL((k_weno<STAGE, CENTRED, false>), L.grid, L.block, 0, ctx->stream, G, R, phi, phin, u, v,
                           out, dt, ghosts);
}

// ------------------------------------------------------------------ run diagnostics
// Per-block partials of: positive cells, sum of their centre x and y, interface length (without
// dx dy), sum |e|, sum e^2, positive cells of the reference, max |e| (e = phi - ref); fixed
// reduction order, so the result does not depend on timing.
constexpr int DG_NV = 8;

__device__ __forceinline__ void dg_block_reduce(double (&a)[DG_NV], double* __restrict__ part) {
    __shared__ double sh[DG_NV][256];
#pragma unroll
    for (int k = 0; k < DG_NV; k++) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
#pragma unroll
            for (int k = 0; k < DG_NV - 1; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
            sh[DG_NV - 1][threadIdx.x] = fmax(sh[DG_NV - 1][threadIdx.x], sh[DG_NV - 1][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < DG_NV; k++) part[k * gridDim.x + blockIdx.x] = sh[k][0];
    }
}

__global__ __launch_bounds__(256) void k_ls_diag(Geom G, const double* __restrict__ ls, const double* __restrict__ ref,
                                                 double bx, double by, double eps, double* __restrict__ part) {
    double a[DG_NV] = {};
    const int64_t N = static_cast<int64_t>(G.nx) * G.ny;
    const double hx = 0.5 / G.dx, hy = 0.5 / G.dy, pe = M_PI / eps, he = 0.5 / eps;
    for (int64_t t = blockIdx.x * 256LL + threadIdx.x; t < N; t += static_cast<int64_t>(gridDim.x) * 256) {
        const int i = G.g + static_cast<int>(t / G.ny), j = G.g + static_cast<int>(t % G.ny);
        const int64_t c = G.idx(i, j);
        const double p = ls[c];
        if (p > 0.0) {
            a[0] += 1.0;
            a[1] += bx + (static_cast<double>(i - G.g) + 0.5) * G.dx;
            a[2] += by + (static_cast<double>(j - G.g) + 0.5) * G.dy;
        }
        if (fabs(p) < eps) {
            const double gx = (ls[G.idx(i + 1, j)] - ls[G.idx(i - 1, j)]) * hx;
            const double gy = (ls[G.idx(i, j + 1)] - ls[G.idx(i, j - 1)]) * hy;
            a[3] += (1.0 + cos(pe * p)) * he * sqrt(gx * gx + gy * gy);
        }
        if (ref) {
            const double r = ref[c], e = fabs(p - r);
            a[4] += e;
            a[5] += e * e;
            if (r > 0.0) a[6] += 1.0;
            a[7] = fmax(a[7], e);
        }
    }
    dg_block_reduce(a, part);
}

// totals of nb partials per value (one block)
__global__ __launch_bounds__(256) void k_ls_diag_final(const double* __restrict__ part, int nb,
                                                       double* __restrict__ out) {
    double a[DG_NV] = {};
    for (int b = threadIdx.x; b < nb; b += 256) {
#pragma unroll
        for (int k = 0; k < DG_NV - 1; k++) a[k] += part[k * nb + b];
        a[DG_NV - 1] = fmax(a[DG_NV - 1], part[(DG_NV - 1) * nb + b]);
    }
    __shared__ double sh[DG_NV][256];
#pragma unroll
    for (int k = 0; k < DG_NV; k++) sh[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
#pragma unroll
            for (int k = 0; k < DG_NV - 1; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
            sh[DG_NV - 1][threadIdx.x] = fmax(sh[DG_NV - 1][threadIdx.x], sh[DG_NV - 1][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < DG_NV) out[threadIdx.x] = sh[threadIdx.x][0];
}

}  // namespace asd2

using namespace asd2;

namespace {

// after a kernel that could not write the ghosts itself (LS-free sides)
int finish_ghosts(asd2_ctx* ctx, bool fused, double* d) {
    if (fused) return ASD2_OK;
    return bc_apply_p(ctx, d);
}

// The ghost rule the stage kernels evaluate themselves; `fused` is false when a side's rule needs
// the whole array (LS-free extrapolation, periodic copies: the kernels then leave the ghosts and the
// P BC runs after them, as ApplyBC_P_2D does in the reference)
int p_rule(asd2_ctx* ctx, GhostRule* R, bool* fused) {
    *R = ghost_rule_p(ctx, fused);  // (grids are at least g cells wide: every mirror is interior)
    return ASD2_OK;
}

double smin_host(double a, double b) { return (b < a) ? b : a; }

bool has_custom(const asd2_ctx* ctx) {
    for (int s = 0; s < 4; s++)
        if (ctx->bc[ASD2_GRID_P][s] == ASD2_BC_CUSTOM) return true;
    return false;
}

bool has_periodic(const asd2_ctx* ctx) {
    for (int s = 0; s < 4; s++)
        if (ctx->bc[ASD2_GRID_P][s] == ASD2_BC_PERIODIC) return true;
    return false;
}

// Substage B and the average of one reinitialisation iteration with the P BC of LS2 in between
// (Reinit_MinRK2_2D, Reinit_Original_2D, FirstTimeOnlyReinit_Sussman_2D: src/levelset2d.cpp:474,
// :299, :419), for grids with a periodic side.  The reference's periodic east fill writes interior
// rows (bc2d.cpp:792-796), so there a BC changes interior values and two BCs differ from one: these
// grids replay the reference's sequence of BCs exactly.  LS2 into ls2 (SUB 0 from LS1, SUB 3 from
// phi), BC(LS2), phi = (phi + LS2) / 2, and with bc_avg the BC of the reference's average buffer
// (MinRK2 and Original; its interior is what they copy back).
template <bool SUBCELL, int SUB>
int reinit_b_periodic(asd2_ctx* ctx, const GhostRule& R, const double* ls1, const double* ls0, double* d_ls,
                      double* ls2, bool bc_avg) {
    const Geom& G = ctx->geo;
    const Launch L = tiles(G, RJ, RI);
    hipLaunchKernelGGL((k_reinit<SUBCELL, SUB>), L.grid, L.block, 0, ctx->stream, G, R, ls1, ls0, d_ls, ls2, 0);
    ASD2_LAUNCH_CHECK();
    ASD2_TRY(bc_apply_p(ctx, ls2));
    hipLaunchKernelGGL(k_ls_average, dim3(cdiv(G.ny, 64), cdiv(G.nx, 4)), dim3(256), 0, ctx->stream, G, d_ls, ls2);
    ASD2_LAUNCH_CHECK();
    return bc_avg ? bc_apply_p(ctx, d_ls) : ASD2_OK;
}

// The three SSP-RK3 stages of Solve_LevelSet_2D when every P side's ghost rule can be fused into
// the stage kernels (no LS-free side): d_ls -> p1 -> p2 -> d_ls, each stage writing its own
// ghosts.  `prof` brackets the launches with the per-launch timing events.
int advect_stages(asd2_ctx* ctx, const GhostRule& R, double* d_ls, const double* d_u, const double* d_v, double dt,
                  double* p1, double* p2, bool prof) {
    const Geom& G = ctx->geo;
    if (prof) prof_begin(ctx, ASD2_KID_WENO_STAGE);
    launch_weno<1, false>(ctx, R, d_ls, d_ls, d_u, d_v, p1, dt, 1);
    if (prof) prof_end(ctx, ASD2_KID_WENO_STAGE);
    ASD2_LAUNCH_CHECK();
    if (prof) prof_begin(ctx, ASD2_KID_WENO_STAGE);
    launch_weno<2, false>(ctx, R, p1, d_ls, d_u, d_v, p2, dt, 1);
    if (prof) prof_end(ctx, ASD2_KID_WENO_STAGE);
    ASD2_LAUNCH_CHECK();
    if (prof) prof_begin(ctx, ASD2_KID_WENO_STAGE);
    launch_weno<3, false>(ctx, R, p2, d_ls, d_u, d_v, d_ls, dt, 1);
    if (prof) prof_end(ctx, ASD2_KID_WENO_STAGE);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

// Stage arrays of asd2_ls_advect_noalloc: one copy per GPU, held until the process ends.  Host
// code must not use this name as an address; it uses the per-device pointer from
// noalloc_scratch().
static __device__ __align__(256) double scratch[2 * static_cast<size_t>(ASD2_LS_NOALLOC_MAX_CELLS)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device, looked up once per device
int noalloc_scratch(double** out) {
    int dev = 0;
    ASD2_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device %d out of range for the static stage buffer", dev);
        return ASD2_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        ASD2_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("hipGetSymbolAddress returned a null stage buffer on device %d", dev);
            return ASD2_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return ASD2_OK;
}

}  // namespace

extern "C" {

int asd2_ls_advect(asd2_ctx* ctx, double* d_ls, const double* d_u, const double* d_v, double dt) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_u || !d_v) { set_error("null array"); return ASD2_ERR_ARG; }
    GhostRule R;
    bool fused;
    ASD2_TRY(p_rule(ctx, &R, &fused));
    const Geom& G = ctx->geo;
    const size_t n = static_cast<size_t>(G.size());
    double *p1, *p2;
    ASD2_TRY(ws(ctx, "ls_stage1", n, &p1));
    ASD2_TRY(ws(ctx, "ls_stage2", n, &p2));
    const bool periodic = has_periodic(ctx);
    if (has_custom(ctx) || periodic) {  // fresh stage arrays: zero ghosts on CUSTOM sides and in the
        ASD2_TRY(zero_async(ctx, p1, n * 8));  // east ghosts a periodic east side never writes
        ASD2_TRY(zero_async(ctx, p2, n * 8));
    }
    ASD2_TRY(bc_apply_p(ctx, d_ls));
    if (!fused) {
        // LS-free sides: build tmpU/tmpV with the full P BC, then run the centred variant
        double *uc, *vc;
        ASD2_TRY(ws(ctx, "ls_uc", n, &uc));
        ASD2_TRY(ws(ctx, "ls_vc", n, &vc));
        ASD2_TRY(zero_async(ctx, uc, n * 8));
        ASD2_TRY(zero_async(ctx, vc, n * 8));
        hipLaunchKernelGGL(k_centre_velocity, dim3(cdiv(G.ny, 64), cdiv(G.nx, 4)), dim3(256), 0, ctx->stream, G,
                           d_u, d_v, uc, vc);
        ASD2_LAUNCH_CHECK();
        ASD2_TRY(bc_apply_p(ctx, uc));
        ASD2_TRY(bc_apply_p(ctx, vc));
        launch_weno<1, true>(ctx, R, d_ls, d_ls, uc, vc, p1, dt, 0);
        ASD2_LAUNCH_CHECK();
        // LS1 and LS2 get the P BC twice in the reference (after the stage and inside the next
        // HJWENO5_LS_2D); with a periodic side the second one changes them again
        ASD2_TRY(bc_apply_p(ctx, p1));
        if (periodic) ASD2_TRY(bc_apply_p(ctx, p1));
        launch_weno<2, true>(ctx, R, p1, d_ls, uc, vc, p2, dt, 0);
        ASD2_LAUNCH_CHECK();
        ASD2_TRY(bc_apply_p(ctx, p2));
        if (periodic) ASD2_TRY(bc_apply_p(ctx, p2));
        launch_weno<3, true>(ctx, R, p2, d_ls, uc, vc, d_ls, dt, 0);
        ASD2_LAUNCH_CHECK();
        return bc_apply_p(ctx, d_ls);
    }
    return advect_stages(ctx, R, d_ls, d_u, d_v, dt, p1, p2, true);
}

int asd2_ls_advect_noalloc(asd2_ctx* ctx, double* d_ls, const double* d_u, const double* d_v, double dt) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_u || !d_v) { set_error("null array"); return ASD2_ERR_ARG; }
    GhostRule R;
    bool fused;
    ASD2_TRY(p_rule(ctx, &R, &fused));
    if (!fused) {
        set_error("asd2_ls_advect_noalloc: level-set-free and periodic pressure sides need two more work "
                  "arrays; use asd2_ls_advect");
        return ASD2_ERR_UNSUPPORTED;
    }
    const Geom& G = ctx->geo;
    const int64_t n = G.size();
    if (n > ASD2_LS_NOALLOC_MAX_CELLS) {
        set_error("asd2_ls_advect_noalloc: %lld padded cells exceed the static stage buffer of %d cells per stage",
                  static_cast<long long>(n), ASD2_LS_NOALLOC_MAX_CELLS);
        return ASD2_ERR_ARG;
    }
    double* base = nullptr;
    ASD2_TRY(noalloc_scratch(&base));
    double* p1 = base;
    double* p2 = base + ASD2_LS_NOALLOC_MAX_CELLS;
    // the buffer keeps the previous call's values: CUSTOM sides read zero ghosts in the stages
    if (has_custom(ctx)) {
        ASD2_TRY(zero_async(ctx, p1, static_cast<size_t>(n) * 8));
        ASD2_TRY(zero_async(ctx, p2, static_cast<size_t>(n) * 8));
    }
    ASD2_TRY(bc_apply_p(ctx, d_ls));
    return advect_stages(ctx, R, d_ls, d_u, d_v, dt, p1, p2, false);
}

int asd2_ls_weno_rhs(asd2_ctx* ctx, double* d_ls, const double* d_uc, const double* d_vc, double* d_L) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_uc || !d_vc || !d_L) { set_error("null array"); return ASD2_ERR_ARG; }
    GhostRule R;
    bool fused;
    ASD2_TRY(p_rule(ctx, &R, &fused));
    const Geom& G = ctx->geo;
    ASD2_TRY(bc_apply_p(ctx, d_ls));
    ASD2_TRY(zero_async(ctx, d_L, static_cast<size_t>(G.size()) * 8));
    launch_weno<0, true>(ctx, R, d_ls, d_ls, d_uc, d_vc, d_L, 0.0, 0);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_ls_subcell_absgrad(asd2_ctx* ctx, const double* d_ls, const double* d_ls0, double* d_out) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_ls0 || !d_out) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    ASD2_TRY(zero_async(ctx, d_out, static_cast<size_t>(G.size()) * 8));
    const Launch L = tiles(G, RJ, RI);
    hipLaunchKernelGGL(k_absgrad<true>, L.grid, L.block, 0, ctx->stream, G, d_ls, d_ls0, d_out);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_ls_eno_absgrad(asd2_ctx* ctx, const double* d_ls, const double* d_ls0, double* d_out) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_ls0 || !d_out) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    ASD2_TRY(zero_async(ctx, d_out, static_cast<size_t>(G.size()) * 8));
    const Launch L = tiles(G, RJ, RI);
    hipLaunchKernelGGL(k_absgrad<false>, L.grid, L.block, 0, ctx->stream, G, d_ls, d_ls0, d_out);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_ls_reinit_iterations(double eps, double max_time) {
    const double adt = eps * 0.5;
    if (!(adt > 0)) return 0;
    double t = 0.0;
    int k = 0;
    while (t < max_time) { t += adt; k++; }
    return k;
}

int asd2_ls_reinit_minrk2(asd2_ctx* ctx, double* d_ls, int iterations) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || iterations < 0) { set_error("bad argument"); return ASD2_ERR_ARG; }
    GhostRule R;
    bool fused;
    ASD2_TRY(p_rule(ctx, &R, &fused));
    const Geom& G = ctx->geo;
    const size_t n = static_cast<size_t>(G.size());
    double *ls0, *ls1;
    ASD2_TRY(ws(ctx, "ri_ls0", n, &ls0));
    ASD2_TRY(ws(ctx, "ri_ls1", n, &ls1));
    if (has_custom(ctx)) ASD2_TRY(zero_async(ctx, ls1, n * 8));
    ASD2_TRY(bc_apply_p(ctx, d_ls));
    ASD2_HIP(hipMemcpyAsync(ls0, d_ls, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    if (fused && !has_custom(ctx)) {
        // one launch per iteration, phi ping-ponged between d_ls and ls1 so that the last
        // iteration writes d_ls; with an odd count the first iteration reads phi from ls0 (equal
        // to d_ls after the BC)
        const Launch F = tiles(G, FJ, RFI, RFT);
        const double* src = (iterations & 1) ? ls0 : d_ls;
        for (int it = 0; it < iterations; it++) {
            double* dst = ((iterations - it) & 1) ? d_ls : ls1;
            prof_begin(ctx, ASD2_KID_REINIT);
            hipLaunchKernelGGL((k_reinit_minrk2<RFI, RFT>), F.grid, F.block, 0, ctx->stream, G, R, src, ls0, dst);
            prof_end(ctx, ASD2_KID_REINIT);
            ASD2_LAUNCH_CHECK();
            src = dst;
        }
        return ASD2_OK;
    }
    const Launch L = tiles(G, RJ, RI);
    const int gh = fused ? 1 : 0;
    // A periodic side replays the reference's BCs one for one (reinit_b_periodic): the BC of phi
    // at the start of every iteration (here the first one; later ones end the previous iteration),
    // BC(LS1), BC(LS2), the BC of the average; LS1 and LS2 start from zero as the reference's
    // fresh vectors do (a periodic east side never writes their east ghosts).
    const bool periodic = has_periodic(ctx);
    double* ls2 = nullptr;
    if (periodic) {
        ASD2_TRY(ws(ctx, "ri_ls2", n, &ls2));
        ASD2_TRY(zero_async(ctx, ls1, n * 8));
        ASD2_TRY(zero_async(ctx, ls2, n * 8));
        ASD2_TRY(bc_apply_p(ctx, d_ls));
    }
    for (int it = 0; it < iterations; it++) {
        prof_begin(ctx, ASD2_KID_REINIT);
        hipLaunchKernelGGL((k_reinit<true, 0>), L.grid, L.block, 0, ctx->stream, G, R, d_ls, ls0, d_ls, ls1, gh);
        prof_end(ctx, ASD2_KID_REINIT);
        ASD2_LAUNCH_CHECK();
        ASD2_TRY(finish_ghosts(ctx, fused, ls1));
        if (periodic) {
            ASD2_TRY((reinit_b_periodic<true, 0>(ctx, R, ls1, ls0, d_ls, ls2, true)));
        } else {
            prof_begin(ctx, ASD2_KID_REINIT);
            hipLaunchKernelGGL((k_reinit<true, 1>), L.grid, L.block, 0, ctx->stream, G, R, ls1, ls0, d_ls, d_ls, gh);
            prof_end(ctx, ASD2_KID_REINIT);
            ASD2_LAUNCH_CHECK();
        }
        ASD2_TRY(finish_ghosts(ctx, fused, d_ls));
    }
    return ASD2_OK;
}

int asd2_ls_reinit_sussman(asd2_ctx* ctx, double* d_ls, int iterations) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || iterations < 0) { set_error("bad argument"); return ASD2_ERR_ARG; }
    GhostRule R;
    bool fused;
    ASD2_TRY(p_rule(ctx, &R, &fused));
    const Geom& G = ctx->geo;
    const size_t n = static_cast<size_t>(G.size());
    double *ls0, *ls1, *tmp;
    ASD2_TRY(ws(ctx, "ri_ls0", n, &ls0));
    ASD2_TRY(ws(ctx, "ri_ls1", n, &ls1));
    ASD2_TRY(ws(ctx, "ri_tmp", n, &tmp));
    // LS1 never gets a BC (its ghosts stay 0) and tmp's corners stay 0 (levelset2d.cpp:326-333)
    ASD2_TRY(zero_async(ctx, ls1, n * 8));
    ASD2_TRY(zero_async(ctx, tmp, n * 8));
    ASD2_HIP(hipMemcpyAsync(ls0, d_ls, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    ASD2_TRY(bc_apply_p(ctx, d_ls));
    const Launch L = tiles(G, RJ, RI);
    const int gh = fused ? 1 : 0;
    for (int it = 0; it < iterations; it++) {
        hipLaunchKernelGGL((k_reinit<false, 0>), L.grid, L.block, 0, ctx->stream, G, R, d_ls, ls0, d_ls, ls1, 0);
        ASD2_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_reinit<false, 1>), L.grid, L.block, 0, ctx->stream, G, R, ls1, ls0, d_ls, tmp, gh);
        ASD2_LAUNCH_CHECK();
        ASD2_TRY(finish_ghosts(ctx, fused, tmp));
        hipLaunchKernelGGL(k_sussman_constraint, dim3(cdiv(G.ny, SJ), cdiv(G.nx, SI)), dim3(WTH), 0, ctx->stream,
                           G, R, tmp, ls0, static_cast<const double*>(nullptr), d_ls, 1, gh);
        ASD2_LAUNCH_CHECK();
        ASD2_TRY(finish_ghosts(ctx, fused, d_ls));
    }
    return ASD2_OK;
}

// Reinit_Original_2D: Sussman's ENO2 substages with BC on LS1 and no constraint; lsInit is
// copied before the first BC (src/levelset2d.cpp:264-316)
int asd2_ls_reinit_original(asd2_ctx* ctx, double* d_ls, int iterations) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || iterations < 0) { set_error("bad argument"); return ASD2_ERR_ARG; }
    GhostRule R;
    bool fused;
    ASD2_TRY(p_rule(ctx, &R, &fused));
    const Geom& G = ctx->geo;
    const size_t n = static_cast<size_t>(G.size());
    double *ls0, *ls1;
    ASD2_TRY(ws(ctx, "ri_ls0", n, &ls0));
    ASD2_TRY(ws(ctx, "ri_ls1", n, &ls1));
    ASD2_TRY(zero_async(ctx, ls1, n * 8));
    ASD2_HIP(hipMemcpyAsync(ls0, d_ls, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    ASD2_TRY(bc_apply_p(ctx, d_ls));
    const Launch L = tiles(G, RJ, RI);
    const int gh = fused ? 1 : 0;
    const bool periodic = has_periodic(ctx);  // BCs replayed one for one (reinit_b_periodic)
    double* ls2 = nullptr;
    if (periodic) {
        ASD2_TRY(ws(ctx, "ri_ls2", n, &ls2));
        ASD2_TRY(zero_async(ctx, ls2, n * 8));
    }
    for (int it = 0; it < iterations; it++) {
        hipLaunchKernelGGL((k_reinit<false, 0>), L.grid, L.block, 0, ctx->stream, G, R, d_ls, ls0, d_ls, ls1, gh);
        ASD2_LAUNCH_CHECK();
        ASD2_TRY(finish_ghosts(ctx, fused, ls1));
        if (periodic) {
            ASD2_TRY((reinit_b_periodic<false, 0>(ctx, R, ls1, ls0, d_ls, ls2, true)));
            ASD2_TRY(finish_ghosts(ctx, fused, d_ls));
            continue;
        }
        hipLaunchKernelGGL((k_reinit<false, 1>), L.grid, L.block, 0, ctx->stream, G, R, ls1, ls0, d_ls, d_ls, gh);
        ASD2_LAUNCH_CHECK();
        ASD2_TRY(finish_ghosts(ctx, fused, d_ls));
    }
    return ASD2_OK;
}

int asd2_ls_reinit_first_time(asd2_ctx* ctx, double* d_ls, int iterations) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || iterations < 0) { set_error("bad argument"); return ASD2_ERR_ARG; }
    GhostRule R;
    bool fused;
    ASD2_TRY(p_rule(ctx, &R, &fused));
    const Geom& G = ctx->geo;
    const size_t n = static_cast<size_t>(G.size());
    double *ls0, *ls1;
    ASD2_TRY(ws(ctx, "ri_ls0", n, &ls0));
    ASD2_TRY(ws(ctx, "ri_ls1", n, &ls1));
    ASD2_TRY(zero_async(ctx, ls1, n * 8));
    // phi0 is taken before any BC; its normalisation by |grad phi0| (:389-396) keeps its sign
    // and its zero set, the only properties the loop reads
    ASD2_HIP(hipMemcpyAsync(ls0, d_ls, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
    ASD2_TRY(bc_apply_p(ctx, d_ls));
    const Launch L = tiles(G, RJ, RI);
    const int gh = fused ? 1 : 0;
    // A periodic side replays the reference's BCs one for one: BC of the normalised phi0 (the
    // normalisation keeps every sign, and only signs are read), BC(LS2) before the average
    // (reinit_b_periodic); LS2 starts from zero as the reference's fresh vector does.
    const bool periodic = has_periodic(ctx);
    double* ls2 = nullptr;
    if (periodic) {
        ASD2_TRY(ws(ctx, "ri_ls2", n, &ls2));
        ASD2_TRY(zero_async(ctx, ls2, n * 8));
        ASD2_TRY(bc_apply_p(ctx, ls0));
    }
    for (int it = 0; it < iterations; it++) {
        hipLaunchKernelGGL((k_reinit<false, 0>), L.grid, L.block, 0, ctx->stream, G, R, d_ls, ls0, d_ls, ls1, gh);
        ASD2_LAUNCH_CHECK();
        ASD2_TRY(finish_ghosts(ctx, fused, ls1));
        if (periodic) {
            ASD2_TRY((reinit_b_periodic<false, 3>(ctx, R, ls1, ls0, d_ls, ls2, false)));
            ASD2_TRY(finish_ghosts(ctx, fused, d_ls));
            continue;
        }
        hipLaunchKernelGGL((k_reinit<false, 2>), L.grid, L.block, 0, ctx->stream, G, R, ls1, ls0, d_ls, d_ls, gh);
        ASD2_LAUNCH_CHECK();
        ASD2_TRY(finish_ghosts(ctx, fused, d_ls));
    }
    return ASD2_OK;
}

int asd2_ls_weno_absgrad(asd2_ctx* ctx, const double* d_ls, const double* d_ls0, double* d_out) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_ls0 || !d_out) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    ASD2_TRY(zero_async(ctx, d_out, static_cast<size_t>(G.size()) * 8));
    hipLaunchKernelGGL(k_weno_absgrad, dim3(cdiv(G.ny, 64), cdiv(G.nx, 4)), dim3(256), 0, ctx->stream, G, d_ls,
                       d_ls0, d_out);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_ls_heaviside_deriv(asd2_ctx* ctx, const double* d_ls, double* d_out) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_out) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    hipLaunchKernelGGL(k_pointwise, dim3(1024), dim3(256), 0, ctx->stream, G.size(), 0, smin_host(G.dx, G.dy),
                       d_ls, d_out);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_ls_smoothed_sign(asd2_ctx* ctx, const double* d_ls, double* d_out) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_out) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    hipLaunchKernelGGL(k_pointwise, dim3(1024), dim3(256), 0, ctx->stream, G.size(), 1, smin_host(G.dx, G.dy),
                       d_ls, d_out);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_ls_sussman_constraint(asd2_ctx* ctx, const double* d_ls, const double* d_ls0, const double* d_hd,
                               double* d_out) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !d_ls0 || !d_hd || !d_out) { set_error("null array"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    ASD2_TRY(zero_async(ctx, d_out, static_cast<size_t>(G.size()) * 8));
    GhostRule R{};
    hipLaunchKernelGGL(k_sussman_constraint, dim3(cdiv(G.ny, SJ), cdiv(G.nx, SI)), dim3(WTH), 0, ctx->stream, G, R,
                       d_ls, d_ls0, d_hd, d_out, 0, 0);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_ls_diagnostics(asd2_ctx* ctx, const double* d_ls, const double* d_ref, asd2_ls_diag* out) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ls || !out) { set_error("null argument"); return ASD2_ERR_ARG; }
    const Geom& G = ctx->geo;
    const int nb = std::max(1, std::min(1024, static_cast<int>((static_cast<int64_t>(G.nx) * G.ny + 255) / 256)));
    double *part, *res;
    ASD2_TRY(ws(ctx, "diag_part", static_cast<int64_t>(DG_NV) * 1024, &part));
    ASD2_TRY(ws(ctx, "diag_out", DG_NV, &res));
    const double eps = 1.5 * smin_host(G.dx, G.dy);
    hipLaunchKernelGGL(k_ls_diag, dim3(nb), dim3(256), 0, ctx->stream, G, d_ls, d_ref, ctx->base_x, ctx->base_y, eps,
                       part);
    hipLaunchKernelGGL(k_ls_diag_final, dim3(1), dim3(256), 0, ctx->stream, part, nb, res);
    ASD2_LAUNCH_CHECK();
    double h[DG_NV];
    ASD2_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    ASD2_HIP(hipStreamSynchronize(ctx->stream));
    asd2_ls_diag d{};
    d.cells = h[0];
    d.mass = h[0] * G.dx * G.dy;
    d.centroid_x = h[0] > 0.0 ? h[1] / h[0] : 0.0;
    d.centroid_y = h[0] > 0.0 ? h[2] / h[0] : 0.0;
    d.length = h[3] * G.dx * G.dy;
    d.circularity = d.length > 0.0 ? 2.0 * std::sqrt(M_PI * d.mass) / d.length : 0.0;
    if (d_ref) {
        d.err_linf = h[7];
        d.err_l1 = h[4];
        d.err_l2 = std::sqrt(h[5]) / (static_cast<double>(G.nx) * G.ny);
        d.ref_cells = h[6];
        d.mass_loss_pct = h[6] > 0.0 ? std::fabs(h[0] - h[6]) / h[6] * 100.0 : 0.0;
    }
    *out = d;
    return ASD2_OK;
}

int asd2_ls_unit_weno5(asd2_ctx* ctx, const double* d_F, double* d_FP, double* d_FM, double d, int n, int g) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_F || !d_FP || !d_FM || n <= 0 || g != 3 || !(d > 0)) { set_error("bad argument"); return ASD2_ERR_ARG; }
    hipLaunchKernelGGL(k_unit_weno5, dim3(cdiv(n, 256)), dim3(256), 0, ctx->stream, d_F, d_FP, d_FM, d, n, g);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

}  // extern "C"
