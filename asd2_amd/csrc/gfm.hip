// gfm.hip -- ghost-fluid assembly of the variable-density pressure Poisson system,
// stored matrix-free (reference MACSolver2D::SolvePoisson, src/mac2d.cpp:1249-1656).
//
// Per interior cell the kernel writes the CSR values the reference would push for the
// cell's row: ax = west neighbour, ay = south neighbour (0 where the reference drops
// the neighbour), diag = centre, rhs.  The east/north neighbours of a row are the
// west/south values of the next cells, bit-identical to the reference's own east/north
// values, so the operator is symmetric.  The arithmetic follows the reference order
// with contraction off.
#include "internal.h"
#include "solver.h"

namespace asd2 {

// heavy-phase fraction thetaH and curvature weight theta of the face between a (west or
// south value) and b (east or north value); west_of_cell = true when the cell is b.
// src/mac2d.cpp:1345-1358 (W/S faces) and :1375-1388 (E/N faces).
__device__ __forceinline__ void face_theta(double a, double b, bool cell_is_b, double& thH, double& th) {
#pragma clang fp contract(off)
    if (a >= 0.0 && b >= 0.0) { thH = 1.0; th = 0.0; }
    else if (a < 0.0 && b < 0.0) { thH = 0.0; th = 0.0; }
    else if (a >= 0.0 && b < 0.0) {
        thH = fabs(a) / (fabs(a) + fabs(b));
        th = cell_is_b ? fabs(a) / (fabs(a) + fabs(b)) : fabs(b) / (fabs(a) + fabs(b));
    } else {
        thH = fabs(b) / (fabs(a) + fabs(b));
        th = cell_is_b ? fabs(a) / (fabs(a) + fabs(b)) : fabs(b) / (fabs(a) + fabs(b));
    }
}

struct GfmArgs {
    double rho_h, rho_l, sigma, dt;
    int bc[4];
    double c[4];
    int pin;
    double base_r;  // axisymmetric: radius of the west wall
};

// the side kinds the assembly treats as Dirichlet (drop the neighbour, move the wall value to the
// right-hand side): WALL is Dirichlet-like in the planar assembly (src/mac2d.cpp:1521-1528) and
// Neumann-like in the axisymmetric one (src/mac2dAxisym.cpp:1790-1862)
template <bool AXI>
__device__ __forceinline__ bool is_dir(int k) {
    return k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET || k == ASD2_BC_OUTLET || k == ASD2_BC_PRESSURE ||
           (!AXI && k == ASD2_BC_WALL);
}
template <bool AXI>
__device__ __forceinline__ bool is_neu(int k) {
    return k == ASD2_BC_NEUMANN || (AXI && (k == ASD2_BC_AXISYM || k == ASD2_BC_WALL));
}

// Planar (AXI = false): face coefficients beta = 1 / rho_eff.  Axisymmetric (AXI = true,
// MACSolver2DAxisym::SolvePoisson, src/mac2dAxisym.cpp:1488-1950, x = r, y = z): r_eff / rho_eff,
// r_eff the W / E face radius where the face lies in one phase, else the radius of the interface
// point r_M theta + r_nb (1 - theta) (clamped at 0 on the W face); r of the cell centre on the S / N
// faces.  Dirichlet walls contribute twice the wall constant there (the reference's
// -ps + 2c with ps zeroed).
template <bool AXI>
__global__ __launch_bounds__(256) void k_gfm_assemble(Geom G, GfmArgs A, const double* __restrict__ ls,
                                                      const double* __restrict__ kap, const double* __restrict__ Hs,
                                                      const double* __restrict__ div, double* __restrict__ ax,
                                                      double* __restrict__ ay, double* __restrict__ dg,
                                                      double* __restrict__ rhs, double* __restrict__ dinv,
                                                      float* __restrict__ ax32, float* __restrict__ ay32,
                                                      float* __restrict__ dg32, float* __restrict__ dinv32) {
#pragma clang fp contract(off)
    const int j = G.g + blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = G.g + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + G.g || j >= G.ny + G.g) return;
    const int64_t c = G.idx(i, j);
    const int64_t cw = G.idx(i - 1, j), ce = G.idx(i + 1, j), cs = G.idx(i, j - 1), cn = G.idx(i, j + 1);
    const double lsW = ls[cw], lsE = ls[ce], lsM = ls[c], lsS = ls[cs], lsN = ls[cn];
    double HM, HW, HE, HS, HN;
    if (Hs) { HM = Hs[c]; HW = Hs[cw]; HE = Hs[ce]; HS = Hs[cs]; HN = Hs[cn]; }
    else {
        HM = lsM >= 0.0 ? 1.0 : 0.0; HW = lsW >= 0.0 ? 1.0 : 0.0; HE = lsE >= 0.0 ? 1.0 : 0.0;
        HS = lsS >= 0.0 ? 1.0 : 0.0; HN = lsN >= 0.0 ? 1.0 : 0.0;
    }
    const double kM = kap[c];
    const double dx = G.dx, dy = G.dy, dt = A.dt, sigma = A.sigma;
    // radii of the axisymmetric faces (src/mac2dAxisym.cpp:1577-1581)
    const double rPW = A.base_r + (i - 0.5 - G.g) * dx, rPM = A.base_r + (i + 0.5 - G.g) * dx;
    const double rPE = A.base_r + (i + 1.5 - G.g) * dx;
    const double rPWHalf = A.base_r + (i - G.g) * dx, rPEHalf = A.base_r + (i + 1.0 - G.g) * dx;
    double thH, th;
    face_theta(lsW, lsM, true, thH, th);
    double iW;
    if (AXI) {
        double rEff = rPWHalf;
        if ((lsW >= 0.0) != (lsM >= 0.0)) rEff = smax(rPM * th + rPW * (1.0 - th), 0.0);
        iW = rEff / (A.rho_h * thH + A.rho_l * (1.0 - thH));
    } else {
        iW = 1.0 / (A.rho_h * thH + A.rho_l * (1.0 - thH));
    }
    double kE = kM * th + kap[cw] * (1.0 - th);
    double FW = -dt * (-sigma * kE) * (HM - HW);
    FW *= iW / (dx * dx);
    face_theta(lsM, lsE, false, thH, th);
    double iE;
    if (AXI) {
        double rEff = rPEHalf;
        if ((lsM >= 0.0) != (lsE >= 0.0)) rEff = rPM * th + rPE * (1.0 - th);
        iE = rEff / (A.rho_h * thH + A.rho_l * (1.0 - thH));
    } else {
        iE = 1.0 / (A.rho_h * thH + A.rho_l * (1.0 - thH));
    }
    kE = kM * th + kap[ce] * (1.0 - th);
    double FE = dt * (-sigma * kE) * (HE - HM);
    FE *= iE / (dx * dx);
    face_theta(lsS, lsM, true, thH, th);
    const double iS = (AXI ? rPM : 1.0) / (A.rho_h * thH + A.rho_l * (1.0 - thH));
    kE = kM * th + kap[cs] * (1.0 - th);
    double FS = -dt * (-sigma * kE) * (HM - HS);
    FS *= iS / (dy * dy);
    face_theta(lsM, lsN, false, thH, th);
    const double iN = (AXI ? rPM : 1.0) / (A.rho_h * thH + A.rho_l * (1.0 - thH));
    kE = kM * th + kap[cn] * (1.0 - th);
    double FN = dt * (-sigma * kE) * (HN - HM);
    FN *= iN / (dy * dy);
    double r = 0.0;
    r -= FW + FE + FS + FN;
    r -= div[c];
    const double nS = -iS, nW = -iW, nE = -iE, nN = -iN;
    double C = -(nW + nE) / (dx * dx) - (nS + nN) / (dy * dy);
    double aW = nW / (dx * dx), aS = nS / (dy * dy);
    // wall value: the constant (planar), -ps + 2c with ps = 0 (axisymmetric)
    auto wall = [&](int s) { return AXI ? 2.0 * A.c[s] : A.c[s]; };
    if (i == G.g) {
        if (is_neu<AXI>(A.bc[0])) { aW = 0.0; C += nW / (dx * dx); }
        else if (is_dir<AXI>(A.bc[0])) { aW = 0.0; r -= nW / (dx * dx) * wall(0); }
    }
    if (i == G.g + G.nx - 1) {
        if (is_neu<AXI>(A.bc[1])) C += nE / (dx * dx);
        else if (is_dir<AXI>(A.bc[1])) r -= nE / (dx * dx) * wall(1);
    }
    if (j == G.g) {
        if (is_neu<AXI>(A.bc[2])) { aS = 0.0; C += nS / (dy * dy); }
        else if (is_dir<AXI>(A.bc[2])) { aS = 0.0; r -= nS / (dy * dy) * wall(2); }
    }
    if (j == G.g + G.ny - 1) {
        if (is_neu<AXI>(A.bc[3])) C += nN / (dy * dy);
        else if (is_dir<AXI>(A.bc[3])) r -= nN / (dy * dy) * wall(3);
    }
    if (A.pin) {
        if (i == G.g + 1 && j == G.g) aW = 0.0;
        else if (i == G.g && j == G.g + 1) aS = 0.0;
        if (i == G.g && j == G.g) r = 0.0;
    }
    ax[c] = aW;
    ay[c] = aS;
    dg[c] = C;
    rhs[c] = r;
    const double di = 1.0 / C;
    dinv[c] = di;
    if (ax32) {
        // the FP32 operator is applied in flux form, sum a (p_nb - p_c) + d p_c, with d the part of
        // the diagonal no stored coupling accounts for: faces to Dirichlet-family walls, the
        // couplings to the pinned cell, and the pinned cell's whole row (exact here, no cancellation)
        double d = 0.0;
        if (i == G.g && is_dir<AXI>(A.bc[0])) d += -nW / (dx * dx);
        if (i == G.g + G.nx - 1 && is_dir<AXI>(A.bc[1])) d += -nE / (dx * dx);
        if (j == G.g && is_dir<AXI>(A.bc[2])) d += -nS / (dy * dy);
        if (j == G.g + G.ny - 1 && is_dir<AXI>(A.bc[3])) d += -nN / (dy * dy);
        if (A.pin) {
            if (i == G.g + 1 && j == G.g) d += -nW / (dx * dx);
            else if (i == G.g && j == G.g + 1) d += -nS / (dy * dy);
            if (i == G.g && j == G.g) d = C;
        }
        ax32[c] = static_cast<float>(aW);
        ay32[c] = static_cast<float>(aS);
        dg32[c] = static_cast<float>(d);
        dinv32[c] = static_cast<float>(di);
    }
}

int system_arrays(asd2_ctx* ctx, System* S, bool want32) {
    const size_t n = static_cast<size_t>(ctx->geo.size());
    ASD2_TRY(ws(ctx, "sys_ax", n, &S->ax));
    ASD2_TRY(ws(ctx, "sys_ay", n, &S->ay));
    ASD2_TRY(ws(ctx, "sys_diag", n, &S->diag));
    ASD2_TRY(ws(ctx, "sys_rhs", n, &S->rhs));
    ASD2_TRY(ws(ctx, "sys_dinv", n, &S->dinv));
    S->ax32 = S->ay32 = S->diag32 = S->dinv32 = nullptr;
    if (want32) {
        ASD2_TRY(ws(ctx, "sys_ax32", n, &S->ax32));
        ASD2_TRY(ws(ctx, "sys_ay32", n, &S->ay32));
        ASD2_TRY(ws(ctx, "sys_diag32", n, &S->diag32));
        ASD2_TRY(ws(ctx, "sys_dinv32", n, &S->dinv32));
    }
    return ASD2_OK;
}

bool all_neumann_family(const asd2_ctx* ctx) {
    for (int s = 0; s < 4; s++) {
        const int k = ctx->bc[ASD2_GRID_P][s];
        if (!(k == ASD2_BC_NEUMANN || k == ASD2_BC_AXISYM || k == ASD2_BC_WALL)) return false;
    }
    return true;
}

// FP32 copies of an FP64 system (for an imported system): couplings, Jacobi inverse and the
// flux-form remainder d = diag - sum |couplings| of each interior cell, taken as 0 when it is
// rounding noise of the FP64 diagonal
__global__ void k_sys_to32(Geom G, const double* ax, const double* ay, const double* dg, const double* di,
                           float* ax32, float* ay32, float* dg32, float* di32) {
    const int j = G.g + blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = G.g + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + G.g || j >= G.ny + G.g) return;
    const int64_t k = G.idx(i, j);
    const double s = -(ax[k] + ax[k + G.sy] + ay[k] + ay[k + 1]);
    double d = dg[k] - s;
    if (fabs(d) <= 1e-12 * fabs(dg[k])) d = 0.0;
    ax32[k] = static_cast<float>(ax[k]);
    ay32[k] = static_cast<float>(ay[k]);
    dg32[k] = static_cast<float>(d);
    di32[k] = static_cast<float>(di[k]);
}

__global__ void k_dinv(Geom G, const double* dg, double* di) {
    const int j = G.g + blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = G.g + blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.nx + G.g || j >= G.ny + G.g) return;
    di[G.idx(i, j)] = 1.0 / dg[G.idx(i, j)];
}

// The FP32 copies of the assembled system: both producers of a system (the assembly kernel and
// the import) write them together with the FP64 arrays, so only the pointers are fetched here.
int ensure_system32(asd2_ctx* ctx, System* S) {
    if (S->ax32) return ASD2_OK;
    if (!ctx->system_ready) {
        set_error("no assembled system");
        return ASD2_ERR_STATE;
    }
    return system_arrays(ctx, S, true);
}

}  // namespace asd2

using namespace asd2;

extern "C" {

namespace {
int gfm_assemble(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_kappa, const double* d_H,
                 const double* d_div, bool axi) {
    ASD2_TRY(check_ctx(ctx));
    if (!fl || !d_ls || !d_kappa || !d_div) { set_error("null argument"); return ASD2_ERR_ARG; }
    for (int s = 0; s < 4; s++) {
        const int k = ctx->bc[ASD2_GRID_P][s];
        if (k == ASD2_BC_PERIODIC || (!axi && k == ASD2_BC_AXISYM)) {
            // mac2d.cpp:1529-1531 (periodic) and the AXISYM rows of :1516-1583 couple the
            // wrong unknowns; the GPU path refuses them.
            set_error("periodic%s pressure boundaries are not supported by the GFM assembly", axi ? "" : "/axisym");
            return ASD2_ERR_UNSUPPORTED;
        }
    }
    System S;
    ASD2_TRY(system_arrays(ctx, &S, true));
    GfmArgs A;
    A.rho_h = fl->rho_h; A.rho_l = fl->rho_l; A.sigma = fl->sigma; A.dt = fl->dt;
    for (int s = 0; s < 4; s++) { A.bc[s] = ctx->bc[ASD2_GRID_P][s]; A.c[s] = ctx->bcc[ASD2_GRID_P][s]; }
    A.pin = all_neumann_family(ctx) ? 1 : 0;
    A.base_r = ctx->base_x;
    const Geom& G = ctx->geo;
    const dim3 grid(cdiv(G.ny, 64), cdiv(G.nx, 4));
    prof_begin(ctx, ASD2_KID_GFM);
    if (axi)
        hipLaunchKernelGGL(k_gfm_assemble<true>, grid, dim3(256), 0, ctx->stream, G, A, d_ls, d_kappa, d_H, d_div,
                           S.ax, S.ay, S.diag, S.rhs, S.dinv, S.ax32, S.ay32, S.diag32, S.dinv32);
    else
        hipLaunchKernelGGL(k_gfm_assemble<false>, grid, dim3(256), 0, ctx->stream, G, A, d_ls, d_kappa, d_H, d_div,
                           S.ax, S.ay, S.diag, S.rhs, S.dinv, S.ax32, S.ay32, S.diag32, S.dinv32);
    prof_end(ctx, ASD2_KID_GFM);
    ASD2_LAUNCH_CHECK();
    ctx->system_ready = true;
    ctx->dres_ring_only = true;
    return ASD2_OK;
}
}  // namespace

int asd2_gfm_assemble(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_kappa,
                      const double* d_H, const double* d_div) {
    return gfm_assemble(ctx, fl, d_ls, d_kappa, d_H, d_div, false);
}

int asd2_gfm_assemble_axisym(asd2_ctx* ctx, const asd2_fluid* fl, const double* d_ls, const double* d_kappa,
                             const double* d_H, const double* d_div4p) {
    return gfm_assemble(ctx, fl, d_ls, d_kappa, d_H, d_div4p, true);
}

int asd2_gfm_export(asd2_ctx* ctx, double* d_ax, double* d_ay, double* d_diag, double* d_rhs) {
    ASD2_TRY(check_ctx(ctx));
    if (!ctx->system_ready) { set_error("no assembled system"); return ASD2_ERR_STATE; }
    System S;
    ASD2_TRY(system_arrays(ctx, &S, false));
    const size_t b = static_cast<size_t>(ctx->geo.size()) * 8;
    if (d_ax) ASD2_HIP(hipMemcpyAsync(d_ax, S.ax, b, hipMemcpyDeviceToDevice, ctx->stream));
    if (d_ay) ASD2_HIP(hipMemcpyAsync(d_ay, S.ay, b, hipMemcpyDeviceToDevice, ctx->stream));
    if (d_diag) ASD2_HIP(hipMemcpyAsync(d_diag, S.diag, b, hipMemcpyDeviceToDevice, ctx->stream));
    if (d_rhs) ASD2_HIP(hipMemcpyAsync(d_rhs, S.rhs, b, hipMemcpyDeviceToDevice, ctx->stream));
    return ASD2_OK;
}

int asd2_gfm_import(asd2_ctx* ctx, const double* d_ax, const double* d_ay, const double* d_diag,
                    const double* d_rhs) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_ax || !d_ay || !d_diag || !d_rhs) { set_error("null argument"); return ASD2_ERR_ARG; }
    System S;
    ASD2_TRY(system_arrays(ctx, &S, true));
    const size_t b = static_cast<size_t>(ctx->geo.size()) * 8;
    ASD2_HIP(hipMemcpyAsync(S.ax, d_ax, b, hipMemcpyDeviceToDevice, ctx->stream));
    ASD2_HIP(hipMemcpyAsync(S.ay, d_ay, b, hipMemcpyDeviceToDevice, ctx->stream));
    ASD2_HIP(hipMemcpyAsync(S.diag, d_diag, b, hipMemcpyDeviceToDevice, ctx->stream));
    ASD2_HIP(hipMemcpyAsync(S.rhs, d_rhs, b, hipMemcpyDeviceToDevice, ctx->stream));
    const Geom& G = ctx->geo;
    hipLaunchKernelGGL(k_dinv, dim3(cdiv(G.ny, 64), cdiv(G.nx, 4)), dim3(256), 0, ctx->stream, G, S.diag, S.dinv);
    ASD2_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sys_to32, dim3(cdiv(G.ny, 64), cdiv(G.nx, 4)), dim3(256), 0, ctx->stream, G, S.ax, S.ay,
                       S.diag, S.dinv, S.ax32, S.ay32, S.diag32, S.dinv32);
    ASD2_LAUNCH_CHECK();
    ctx->system_ready = true;
    ctx->dres_ring_only = false;
    return ASD2_OK;
}

}  // extern "C"
