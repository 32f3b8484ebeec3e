// fastpoisson.hip -- constant-coefficient fast Poisson / Helmholtz solve of
// PoissonSolver2D::MKL_2FUniform_2D (reference src/poisson2d.cpp:90-228).
//
// The reference hands the solve to the MKL Poisson Library: -u_xx - u_yy + q u = f with the
// five-point difference on the node mesh that spans the padded cell centres from the first ghost
// to the last (ax = -dx/2, bx = lenX + dx/2, nx + 1 intervals; node (i, j) is padded cell
// (i + g - 1, j + g - 1)).  A Dirichlet side holds given values on its nodes; a Neumann side's
// nodes are unknowns closed by the central difference across the side (the node outside mirrors
// the node inside: row factor 2 on the inner neighbour).
//
// The solve is separable.  Along y (the contiguous index) the 1D operator T_y has closed-form
// eigenvectors: cos(pi k j / M) for Neumann-Neumann, sin(pi k j / M) for Dirichlet-Dirichlet,
// sin / cos(pi (k - 1/2) j / M) for Dirichlet-Neumann / Neumann-Dirichlet, with eigenvalues
// 4 sin^2(theta_k / 2) / hy^2.  T_y = W^-1 K (W = 1/2 on Neumann nodes, K symmetric), so the
// inverse of the eigenvector matrix is N^-1 V^T W.  Forward and inverse transforms are two FP64
// matrix products (rocBLAS DGEMM, O(mx my^2) flops); in between, one tridiagonal solve along x per
// y mode (Thomas, one thread per mode, rows read coalesced across modes).
//
// Degenerate problem (q = 0, Neumann on every side): the system is solved in its compatible
// projection and the solution returned with zero weighted mean (weights w_x(i) w_y(j)); MKL's own
// additive constant is undocumented.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>

#include "internal.h"

namespace asd2 {
namespace {

constexpr int FB = 256;

// one axis of the node mesh: nodes 0..M, unknowns lo..hi-1
struct Axis {
    int M;         // intervals (n + 1)
    int lo_neu;    // 1: Neumann at node 0 (node 0 is an unknown)
    int hi_neu;    // 1: Neumann at node M
    int first;     // first unknown node
    int count;     // unknowns
    double h;
};

Axis make_axis(int n, int lo_kind, int hi_kind, double len, double d) {
    Axis a{};
    a.M = n + 1;
    a.lo_neu = lo_kind == ASD2_BC_NEUMANN;
    a.hi_neu = hi_kind == ASD2_BC_NEUMANN;
    a.first = a.lo_neu ? 0 : 1;
    const int last = a.hi_neu ? a.M : a.M - 1;
    a.count = last - a.first + 1;
    // MKL's spacing from the reference's rectangle: (bx - ax) / intervals
    const double ax = -0.5 * d, bx = len + 0.5 * d;
    a.h = (bx - ax) / a.M;
    return a;
}

// eigenvector k_idx at node j of the 1D operator on axis A, its weight-norm N_k and theta_k
__device__ __forceinline__ void mode_value(const Axis& A, int k_idx, int j, double* v, double* norm, double* theta) {
    const int64_t M = A.M;
    if (A.lo_neu && A.hi_neu) {  // cos(pi k j / M), k = 0..M
        const int64_t k = k_idx;
        *v = cospi(static_cast<double>((k * j) % (2 * M)) / static_cast<double>(M));
        *norm = (k == 0 || k == M) ? static_cast<double>(M) : 0.5 * static_cast<double>(M);
        *theta = static_cast<double>(k) / static_cast<double>(M);
    } else if (!A.lo_neu && !A.hi_neu) {  // sin(pi k j / M), k = 1..M-1
        const int64_t k = k_idx + 1;
        *v = sinpi(static_cast<double>((k * j) % (2 * M)) / static_cast<double>(M));
        *norm = 0.5 * static_cast<double>(M);
        *theta = static_cast<double>(k) / static_cast<double>(M);
    } else {  // (k - 1/2): sin for Dirichlet at node 0, cos for Neumann at node 0, k = 1..M
        const int64_t k2 = 2 * (k_idx + 1) - 1;
        const double a = static_cast<double>((k2 * j) % (4 * M)) / static_cast<double>(2 * M);
        *v = A.lo_neu ? cospi(a) : sinpi(a);
        *norm = 0.5 * static_cast<double>(M);
        *theta = static_cast<double>(k2) / static_cast<double>(2 * M);
    }
}

// V (column-major: row = unknown node, column = mode), Vinv (column-major: row = mode, column =
// node) and the eigenvalues
__global__ __launch_bounds__(FB) void k_fp_basis(Axis A, double* __restrict__ V, double* __restrict__ Vinv,
                                                 double* __restrict__ lam) {
    const int64_t m = A.count;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(FB) + threadIdx.x; e < m * m;
         e += static_cast<int64_t>(gridDim.x) * FB) {
        const int k = static_cast<int>(e / m), r = static_cast<int>(e - static_cast<int64_t>(k) * m);
        const int j = A.first + r;
        double v, nk, th;
        mode_value(A, k, j, &v, &nk, &th);
        const double w = ((j == 0 && A.lo_neu) || (j == A.M && A.hi_neu)) ? 0.5 : 1.0;
        V[r + m * k] = v;
        Vinv[k + m * r] = w * v / nk;
        if (r == 0) {
            const double s = sinpi(0.5 * th);
            lam[k] = 4.0 * s * s / (A.h * A.h);
        }
    }
}

struct Sides {
    int kind[4];     // W, E, S, N
    double c[4];
};

// bd_ax[j], bd_bx[j] (j = 0..ny+1), bd_ay[i], bd_by[i] (i = 0..nx+1) as the reference fills them
// from the caller's ps (src/poisson2d.cpp:150-180): 2c - ps at (g, j+g), (nx+g, j+g), (i+g, g),
// (i+g, ny+g); 0 on Neumann sides
__global__ __launch_bounds__(FB) void k_fp_bounds(Geom G, Sides S, const double* __restrict__ ps,
                                                  double* __restrict__ bd) {
    const int nx = G.nx, ny = G.ny, g = G.g;
    const int total = 2 * (ny + 2) + 2 * (nx + 2);
    for (int e = blockIdx.x * FB + threadIdx.x; e < total; e += gridDim.x * FB) {
        double v = 0.0;
        if (e < ny + 2) {
            if (S.kind[0] == ASD2_BC_DIRICHLET) v = 2.0 * S.c[0] - ps[G.idx(g, e + g)];
        } else if (e < 2 * (ny + 2)) {
            const int j = e - (ny + 2);
            if (S.kind[1] == ASD2_BC_DIRICHLET) v = 2.0 * S.c[1] - ps[G.idx(nx + g, j + g)];
        } else if (e < 2 * (ny + 2) + nx + 2) {
            const int i = e - 2 * (ny + 2);
            if (S.kind[2] == ASD2_BC_DIRICHLET) v = 2.0 * S.c[2] - ps[G.idx(i + g, g)];
        } else {
            const int i = e - 2 * (ny + 2) - (nx + 2);
            if (S.kind[3] == ASD2_BC_DIRICHLET) v = 2.0 * S.c[3] - ps[G.idx(i + g, ny + g)];
        }
        bd[e] = v;
    }
}

// F[r][s] (row-major, r = x unknown, s = y unknown): rhs at the node plus the known Dirichlet
// neighbours moved to the right-hand side
__global__ __launch_bounds__(FB) void k_fp_gather(Geom G, Axis X, Axis Y, const double* __restrict__ rhs,
                                                  const double* __restrict__ bd, double* __restrict__ F) {
    const int64_t mx = X.count, my = Y.count;
    const int nx = G.nx, ny = G.ny, g = G.g;
    const double* bax = bd;
    const double* bbx = bd + (ny + 2);
    const double* bay = bd + 2 * (ny + 2);
    const double* bby = bd + 2 * (ny + 2) + (nx + 2);
    const double ihx2 = 1.0 / (X.h * X.h), ihy2 = 1.0 / (Y.h * Y.h);
    for (int64_t e = blockIdx.x * static_cast<int64_t>(FB) + threadIdx.x; e < mx * my;
         e += static_cast<int64_t>(gridDim.x) * FB) {
        const int r = static_cast<int>(e / my), s = static_cast<int>(e - static_cast<int64_t>(r) * my);
        const int i = X.first + r, j = Y.first + s;
        double f = rhs[G.idx(i + g - 1, j + g - 1)];
        if (!X.lo_neu && i == 1) f += bax[j] * ihx2;
        if (!X.hi_neu && i == X.M - 1) f += bbx[j] * ihx2;
        if (!Y.lo_neu && j == 1) f += bay[i] * ihy2;
        if (!Y.hi_neu && j == Y.M - 1) f += bby[i] * ihy2;
        F[e] = f;
    }
}

// (T_x + lam_k + q) u = F[:, k] for every mode k, in place in F; cp is scratch of F's size.
// The mode that is singular (Neumann on both x sides, lam_k + q = 0) is solved in its compatible
// projection with u at the first node 0, then shifted to zero weighted mean.
__global__ __launch_bounds__(FB) void k_fp_tridiag(Axis X, int my, const double* __restrict__ lam, double q,
                                                   double* __restrict__ F, double* __restrict__ cp) {
    const int k = blockIdx.x * FB + threadIdx.x;
    if (k >= my) return;
    const int mx = X.count;
    const double a = 1.0 / (X.h * X.h), s = lam[k] + q;
    const double diag = 2.0 * a + s;
    const bool singular = X.lo_neu && X.hi_neu && s == 0.0;
    auto at = [&](int r) -> int64_t { return static_cast<int64_t>(r) * my + k; };
    auto sub = [&](int r) { return (X.hi_neu && r == mx - 1) ? -2.0 * a : -a; };
    auto sup = [&](int r) { return (X.lo_neu && r == 0) ? -2.0 * a : -a; };
    int r0 = 0;
    if (singular) {
        // project out the weighted mean (weights 1/2 on the Neumann end nodes); u_0 = 0
        double sw = 0.0, sf = 0.0;
        for (int r = 0; r < mx; r++) {
            const double w = (r == 0 || r == mx - 1) ? 0.5 : 1.0;
            sw += w;
            sf += w * F[at(r)];
        }
        const double mean = sf / sw;
        for (int r = 0; r < mx; r++) F[at(r)] -= mean;
        F[at(0)] = 0.0;
        r0 = 1;
    }
    if (mx - r0 <= 0) return;
    // forward sweep
    double c_prev = sup(r0) / diag;
    double d_prev = F[at(r0)] / diag;
    cp[at(r0)] = c_prev;
    F[at(r0)] = d_prev;
    for (int r = r0 + 1; r < mx; r++) {
        const double b = sub(r);
        const double den = diag - b * c_prev;
        const double c = (r < mx - 1) ? sup(r) / den : 0.0;
        const double d = (F[at(r)] - b * d_prev) / den;
        cp[at(r)] = c;
        F[at(r)] = d;
        c_prev = c;
        d_prev = d;
    }
    // back substitution
    double u_next = F[at(mx - 1)];
    for (int r = mx - 2; r >= r0; r--) {
        const double u = F[at(r)] - cp[at(r)] * u_next;
        F[at(r)] = u;
        u_next = u;
    }
    if (singular) {
        double sw = 0.0, su = 0.0;
        for (int r = 0; r < mx; r++) {
            const double w = (r == 0 || r == mx - 1) ? 0.5 : 1.0;
            sw += w;
            su += w * F[at(r)];
        }
        const double mean = su / sw;
        for (int r = 0; r < mx; r++) F[at(r)] -= mean;
    }
}

// the solution at every node into ps: unknowns from U, Dirichlet nodes their values (x sides
// hold the corners)
__global__ __launch_bounds__(FB) void k_fp_scatter(Geom G, Axis X, Axis Y, const double* __restrict__ U,
                                                   const double* __restrict__ bd, double* __restrict__ ps) {
    const int nx = G.nx, ny = G.ny, g = G.g;
    const int64_t nn = static_cast<int64_t>(nx + 2) * (ny + 2);
    const double* bax = bd;
    const double* bbx = bd + (ny + 2);
    const double* bay = bd + 2 * (ny + 2);
    const double* bby = bd + 2 * (ny + 2) + (nx + 2);
    for (int64_t e = blockIdx.x * static_cast<int64_t>(FB) + threadIdx.x; e < nn;
         e += static_cast<int64_t>(gridDim.x) * FB) {
        const int i = static_cast<int>(e / (ny + 2)), j = static_cast<int>(e - static_cast<int64_t>(i) * (ny + 2));
        const int r = i - X.first, s = j - Y.first;
        double v;
        if (r >= 0 && r < X.count && s >= 0 && s < Y.count) v = U[static_cast<int64_t>(r) * Y.count + s];
        else if (i == 0 && !X.lo_neu) v = bax[j];
        else if (i == X.M && !X.hi_neu) v = bbx[j];
        else if (j == 0 && !Y.lo_neu) v = bay[i];
        else v = bby[i];
        ps[G.idx(i + g - 1, j + g - 1)] = v;
    }
}

unsigned grid_for(int64_t work) {
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(8192, (work + FB - 1) / FB)));
}

const char* blas_status(rocblas_status s) { return rocblas_status_to_string(s); }

}  // namespace

void fast_poisson_release(asd2_ctx* ctx) {
    if (ctx && ctx->blas) {
        (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->blas));
        ctx->blas = nullptr;
    }
}

}  // namespace asd2

using namespace asd2;

int asd2_fast_poisson(asd2_ctx* ctx, const double* d_rhs, double* d_ps, double len_x, double len_y, double dx,
                      double dy, double q) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_rhs || !d_ps || !(len_x > 0.0) || !(len_y > 0.0) || !(dx > 0.0) || !(dy > 0.0) || !(q >= 0.0) ||
        !std::isfinite(q) || !std::isfinite(len_x) || !std::isfinite(len_y) || !std::isfinite(dx) ||
        !std::isfinite(dy)) {
        set_error("asd2_fast_poisson: bad argument (rhs, ps, lengths and spacings > 0, finite q >= 0)");
        return ASD2_ERR_ARG;
    }
    const Geom& G = ctx->geo;
    if (G.g < 2) {
        set_error("asd2_fast_poisson: the node mesh needs at least 2 ghost layers (have %d)", G.g);
        return ASD2_ERR_ARG;
    }
    Sides S{};
    for (int s = 0; s < 4; s++) {
        const int k = ctx->bc[0][s];
        if (k != ASD2_BC_NEUMANN && k != ASD2_BC_DIRICHLET) {
            set_error("asd2_fast_poisson: the fast solver takes Neumann or Dirichlet pressure sides only (side %d "
                      "has kind %d)", s, k);
            return ASD2_ERR_UNSUPPORTED;
        }
        S.kind[s] = k;
        S.c[s] = ctx->bcc[0][s];
    }
    const Axis X = make_axis(G.nx, S.kind[0], S.kind[1], len_x, dx);
    const Axis Y = make_axis(G.ny, S.kind[2], S.kind[3], len_y, dy);
    const int64_t mx = X.count, my = Y.count;
    hipStream_t st = ctx->stream;

    // eigenbasis of the y operator, cached per (sides, size, spacing)
    char key[160];
    snprintf(key, sizeof key, "%d %d %d %.17g", Y.M, Y.lo_neu, Y.hi_neu, Y.h);
    double *V = nullptr, *Vinv = nullptr, *lam = nullptr;
    ASD2_TRY(ws(ctx, "fp_V", static_cast<size_t>(my * my), &V));
    ASD2_TRY(ws(ctx, "fp_Vinv", static_cast<size_t>(my * my), &Vinv));
    ASD2_TRY(ws(ctx, "fp_lam", static_cast<size_t>(my), &lam));
    if (ctx->fp_key != key) {
        hipLaunchKernelGGL(k_fp_basis, dim3(grid_for(my * my)), dim3(FB), 0, st, Y, V, Vinv, lam);
        ASD2_LAUNCH_CHECK();
        ctx->fp_key = key;
    }
    double *F = nullptr, *Fh = nullptr, *cp = nullptr, *bd = nullptr;
    ASD2_TRY(ws(ctx, "fp_F", static_cast<size_t>(mx * my), &F));
    ASD2_TRY(ws(ctx, "fp_Fh", static_cast<size_t>(mx * my), &Fh));
    ASD2_TRY(ws(ctx, "fp_cp", static_cast<size_t>(mx * my), &cp));
    ASD2_TRY(ws(ctx, "fp_bd", static_cast<size_t>(2 * (G.nx + 2) + 2 * (G.ny + 2)), &bd));
    if (!ctx->blas) {
        rocblas_handle h = nullptr;
        const rocblas_status rs = rocblas_create_handle(&h);
        if (rs != rocblas_status_success) {
            set_error("asd2_fast_poisson: rocblas_create_handle: %s", blas_status(rs));
            return ASD2_ERR_HIP;
        }
        ctx->blas = h;
    }
    rocblas_handle h = static_cast<rocblas_handle>(ctx->blas);
    rocblas_status rs = rocblas_set_stream(h, st);
    if (rs != rocblas_status_success) {
        set_error("asd2_fast_poisson: rocblas_set_stream: %s", blas_status(rs));
        return ASD2_ERR_HIP;
    }

    prof_begin(ctx, ASD2_KID_FAST_PASS);
    hipLaunchKernelGGL(k_fp_bounds, dim3(grid_for(2 * (G.nx + 2) + 2 * (G.ny + 2))), dim3(FB), 0, st, G, S, d_ps, bd);
    hipLaunchKernelGGL(k_fp_gather, dim3(grid_for(mx * my)), dim3(FB), 0, st, G, X, Y, d_rhs, bd, F);
    prof_end(ctx, ASD2_KID_FAST_PASS);
    ASD2_LAUNCH_CHECK();
    // forward transform: Fh (col-major my x mx) = Vinv (my x my) F (my x mx)
    const double one = 1.0, zero = 0.0;
    const rocblas_int m = static_cast<rocblas_int>(my), n = static_cast<rocblas_int>(mx);
    prof_begin(ctx, ASD2_KID_FAST_GEMM);
    rs = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, m, &one, Vinv, m, F, m, &zero, Fh, m);
    prof_end(ctx, ASD2_KID_FAST_GEMM);
    if (rs != rocblas_status_success) {
        set_error("asd2_fast_poisson: rocblas_dgemm (forward): %s", blas_status(rs));
        return ASD2_ERR_HIP;
    }
    prof_begin(ctx, ASD2_KID_FAST_TRIDIAG);
    hipLaunchKernelGGL(k_fp_tridiag, dim3(static_cast<unsigned>((my + FB - 1) / FB)), dim3(FB), 0, st, X,
                       static_cast<int>(my), lam, q, Fh, cp);
    prof_end(ctx, ASD2_KID_FAST_TRIDIAG);
    ASD2_LAUNCH_CHECK();
    // inverse transform: F = V Fh
    prof_begin(ctx, ASD2_KID_FAST_GEMM);
    rs = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, m, &one, V, m, Fh, m, &zero, F, m);
    prof_end(ctx, ASD2_KID_FAST_GEMM);
    if (rs != rocblas_status_success) {
        set_error("asd2_fast_poisson: rocblas_dgemm (inverse): %s", blas_status(rs));
        return ASD2_ERR_HIP;
    }
    prof_begin(ctx, ASD2_KID_FAST_PASS);
    hipLaunchKernelGGL(k_fp_scatter, dim3(grid_for(static_cast<int64_t>(G.nx + 2) * (G.ny + 2))), dim3(FB), 0, st, G,
                       X, Y, F, bd, d_ps);
    prof_end(ctx, ASD2_KID_FAST_PASS);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}
