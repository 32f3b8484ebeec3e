// krylov.hip -- Jacobi-preconditioned Krylov solvers for the matrix-free GFM system
// (reference PoissonSolver2D, src/poisson2d.cpp:230-575) on gfx950.
//
// Each CG iteration is two streaming kernels: (1) direction update p = z + beta p fused
// with the 5-point SpMV q = A p and the block partials of p.q, (2) the x/r/z update fused
// with the block partials of r.z and r.r.  Each kernel starts by summing the partials of
// the kernel before it: every block reads the same <= 1024 partials in the same order, so
// all blocks derive bit-identical scalars (alpha, beta, the convergence decision), and
// block 0 records them in the device-resident state, into slots indexed by the iteration's
// parity so that no block reads a slot written in the same launch.  A grid-wide "last
// block" reduction would need a device-scope release fence per block, which on this chip
// writes back the XCD's L2 each time (measured: 124 us instead of 16 us for an axpy-dot
// over 4096 blocks); a separate one-block reduction kernel costs about 4 us per launch.
// BiCGStab and the rare bookkeeping steps still use such one-block "tail" kernels.
// Reducing kernels use at most MAX_BLOCKS blocks; each block walks a contiguous range of
// tiles, and consecutive ranges are placed on the same XCD (workgroup b runs on XCD b % 8)
// so stencil halos shared by neighbouring tiles hit the same L2.  The host enqueues
// iterations in batches and polls a pinned copy of the scalars one batch behind; kernels
// of a finished solve exit on the `done` flag.
#include <algorithm>

#include "solver.h"

namespace asd2 {

constexpr int KB = 256;            // threads per block of every kernel here
constexpr int MAX_BLOCKS = 1024;   // blocks of a reducing kernel (= partials per value)
constexpr int TJ = 64, TI = 16;    // SpMV tile (j fastest)
constexpr int NXCD = 8;

__host__ __device__ inline int ceil_div(int64_t a, int64_t b) { return static_cast<int>((a + b - 1) / b); }

// ------------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// block sum of NV values; the totals land in thread 0
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[NV][KB / 64];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) sh[k][w] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0.0;
#pragma unroll
            for (int ww = 0; ww < KB / 64; ww++) s += sh[k][ww];
            v[k] = s;
        }
    }
    __syncthreads();
}

// block partials of a 1-D grid: partials[k * gridDim.x + blockIdx.x]
template <int NV>
__device__ __forceinline__ void put_partials(double (&v)[NV], double* partials) {
    block_sum(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) partials[k * gridDim.x + blockIdx.x] = v[k];
    }
}

// in a one-block tail kernel: totals of nb partials per value, in thread 0
template <int NV>
__device__ __forceinline__ void sum_partials(const double* partials, int nb, double (&v)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += KB) s += partials[k * nb + b];
        v[k] = s;
    }
    block_sum(v);
}

// totals of nb partials per value, in every thread; the summation order is fixed, so every
// block of a launch obtains the same bits
template <int NV>
__device__ __forceinline__ void sum_partials_all(const double* partials, int nb, double (&v)[NV]) {
    __shared__ double bc[NV];
    sum_partials(partials, nb, v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) bc[k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = bc[k];
}

// logical block index: consecutive logical indices run on the same XCD
__device__ __forceinline__ int logical_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    if (nb % NXCD) return b;
    return (b % NXCD) * (nb / NXCD) + b / NXCD;
}

// the contiguous tile range [t0, t1) of this block
__device__ __forceinline__ void tile_range(int ntiles, int& t0, int& t1) {
    const int per = ceil_div(ntiles, gridDim.x);
    t0 = logical_block() * per;
    t1 = min(ntiles, t0 + per);
}

// Visits every interior cell once: 64 x 4 tiles, stacked along i within a block's range.
template <class F>
__device__ __forceinline__ void for_interior(const Geom& G, F f) {
    const int tiles_i = ceil_div(G.nx, 4), ntiles = tiles_i * ceil_div(G.ny, 64);
    int t0, t1;
    tile_range(ntiles, t0, t1);
    for (int t = t0; t < t1; t++) {
        const int tj = t / tiles_i, ti = t - tj * tiles_i;
        const int j = G.g + tj * 64 + (threadIdx.x & 63);
        const int i = G.g + ti * 4 + (threadIdx.x >> 6);
        if (i < G.nx + G.g && j < G.ny + G.g) f(i, j, G.idx(i, j));
    }
}

#define FLAT_LOOP(k, n) \
    for (int64_t k = blockIdx.x * static_cast<int64_t>(KB) + threadIdx.x; k < (n); k += static_cast<int64_t>(gridDim.x) * KB)

// ------------------------------------------------------------------ tail kernels
enum TailOp {
    T_CG_INIT,          // r.z, r.r, b.b  -> bnorm, initial convergence
    T_TRUE_RES,         // ||b - A x||^2  -> reliable update bookkeeping
    T_TRUE_RES_REPORT,  //                -> reported true residual
    T_MX_RESUME,        // (no sum)       -> back to RUNNING after a reliable update
    T_BICG_INIT,        // b.b            -> bnorm, rho
    T_BICG_ALPHA,       // rt.v           -> alpha
    T_BICG_S,           // s.s            -> early exit on ||s||
    T_BICG_OMEGA,       // t.s, t.t       -> omega
    T_BICG_X,           // r.r, rt.r      -> iter++, beta, convergence
};

template <int OP> struct TailNV { static constexpr int value = 1; };
template <> struct TailNV<T_CG_INIT> { static constexpr int value = 3; };
template <> struct TailNV<T_BICG_OMEGA> { static constexpr int value = 2; };
template <> struct TailNV<T_BICG_X> { static constexpr int value = 2; };

// whether the producer of OP's partials ran (the tail mirrors its guard)
template <int OP>
__device__ __forceinline__ bool tail_active(const KrylovState* st) {
    if (OP == T_CG_INIT || OP == T_BICG_INIT || OP == T_TRUE_RES_REPORT) return true;
    if (OP == T_TRUE_RES || OP == T_MX_RESUME) return st->done == DONE_PAUSED;
    if (OP == T_BICG_OMEGA) return st->done == DONE_RUNNING && !(st->flags & 1);
    return st->done == DONE_RUNNING;
}

template <int OP>
__global__ __launch_bounds__(KB) void k_tail(int nb, const double* __restrict__ partials, KrylovState* st) {
    if (!tail_active<OP>(st)) return;
    if (OP == T_MX_RESUME) {
        if (threadIdx.x == 0) st->done = DONE_RUNNING;
        return;
    }
    constexpr int NV = TailNV<OP>::value;
    double v[NV];
    sum_partials(partials, nb, v);
    if (threadIdx.x != 0) return;
    if (OP == T_CG_INIT) {
        double bn = sqrt(v[2]);
        if (bn == 0.0) bn = 1.0;
        st->bnorm = bn;
        st->rr = v[1];
        st->rmax = sqrt(v[1]);
        st->iter = 0;
        st->updates = 0;
        st->fallbacks = 0;
        st->flags = 0;
        st->done = (sqrt(v[1]) / bn <= st->tol) ? DONE_CONVERGED : DONE_RUNNING;
    } else if (OP == T_TRUE_RES) {
        const double rn = sqrt(v[0]);
        st->aux0 = rn;
        st->rmax = rn;
        st->updates += 1;
        if (rn / st->bnorm <= st->tol) st->done = DONE_CONVERGED;
        else if (st->iter >= st->maxit) st->done = DONE_MAXIT;
    } else if (OP == T_TRUE_RES_REPORT) {
        st->aux2 = sqrt(v[0]) / st->bnorm;
    } else if (OP == T_BICG_INIT) {
        double bn = sqrt(v[0]);
        const double rn = bn;
        if (bn == 0.0) bn = 1.0;
        st->bnorm = bn;
        st->rr = v[0];
        st->rho = v[0];  // rtilde . r with rtilde = r = b
        st->alpha = 1.0;
        st->omega = 1.0;
        st->iter = 0;
        st->flags = 0;
        st->done = (rn / bn <= st->tol || v[0] == 0.0) ? DONE_CONVERGED : DONE_RUNNING;
    } else if (OP == T_BICG_ALPHA) {
        st->alpha = st->rho / v[0];
        if (!isfinite(st->alpha)) st->done = DONE_BREAKDOWN;
    } else if (OP == T_BICG_S) {
        st->aux1 = sqrt(v[0]);
        if (st->aux1 / st->bnorm <= st->tol) {
            st->flags |= 1;
            st->omega = 0.0;
        }
    } else if (OP == T_BICG_OMEGA) {
        st->omega = v[0] / v[1];
        if (!isfinite(st->omega)) st->done = DONE_BREAKDOWN;
    } else if (OP == T_BICG_X) {
        const bool half = (st->flags & 1) != 0;
        st->iter += 1;
        st->rr = v[0];
        const double rho_new = v[1];
        const double resid = sqrt(v[0]) / st->bnorm;
        if (half || resid <= st->tol) st->done = DONE_CONVERGED;
        else if (!isfinite(resid)) st->done = DONE_BREAKDOWN;
        else if (rho_new == 0.0) st->done = DONE_BREAKDOWN;
        else if (st->iter >= st->maxit) st->done = DONE_MAXIT;
        st->beta = (rho_new / st->rho) * (st->alpha / st->omega);
        st->rho = rho_new;
    }
}

// ------------------------------------------------------------------ CG kernels
template <typename T>
__global__ __launch_bounds__(KB) void k_cg_init(int64_t n, const double* __restrict__ b, T* __restrict__ x,
                                                T* __restrict__ r, double* __restrict__ r64, double* __restrict__ x64,
                                                const T* __restrict__ dinv, T* __restrict__ z, T* __restrict__ p,
                                                double* partials) {
    double bb = 0.0, rz = 0.0, rr = 0.0;
    FLAT_LOOP(k, n) {
        const double bv = b[k];
        const T rv = static_cast<T>(bv);
        const T zv = dinv[k] * rv;
        x[k] = T(0);
        r[k] = rv;
        z[k] = zv;
        p[k] = rv;
        if (r64) r64[k] = bv;
        if (x64) x64[k] = 0.0;
        bb += bv * bv;
        rz += static_cast<double>(rv) * static_cast<double>(zv);
        rr += static_cast<double>(rv) * static_cast<double>(rv);
    }
    double v[3] = {rz, rr, bb};
    put_partials(v, partials);
}

// The CG recurrence at the start of iteration `it` (poisson2d.cpp:345-395), from the
// partials (r.z, r.r) of the kernel that produced r_it: the convergence, pause (mixed
// precision) or iteration-limit decision and beta.  The reference's iteration 0 uses
// p = rho r + z (beta = rho / 1 with p0 = r0, :331-360) and its iteration 1 restarts with
// p = z (:350-352).  Every block computes the same values; block 0 records them.  Returns
// false when the iteration must not run.
template <bool MIXED>
__device__ __forceinline__ bool cg_begin(KrylovState* st, const double* pu, int nbu, int it, double& beta) {
    double v[2];
    sum_partials_all(pu, nbu, v);
    const double rz = v[0], rr = v[1];
    const double rn = sqrt(rr), resid = rn / st->bnorm;
    if (st->mode == ASD2_CG_REFERENCE) beta = (it == 0) ? rz / 1.0 : (it == 1 ? 0.0 : rz / st->rz[(it - 1) & 1]);
    else beta = (it == 0) ? 0.0 : rz / st->rz[(it - 1) & 1];
    int done = DONE_RUNNING;
    if (!isfinite(resid) || !isfinite(beta)) done = DONE_BREAKDOWN;
    else if (MIXED && (resid <= st->tol || rn < st->delta * st->rmax)) done = DONE_PAUSED;
    else if (!MIXED && resid <= st->tol) done = DONE_CONVERGED;
    else if (it >= st->maxit) done = MIXED ? DONE_PAUSED : DONE_MAXIT;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rr = rr;
        st->iter = it;
        if (done == DONE_RUNNING) {
            st->rz[it & 1] = rz;
            st->beta = beta;
        } else {
            st->done = done;
        }
    }
    return done == DONE_RUNNING;
}

// alpha of iteration `it` from the p.q partials; false (and BREAKDOWN recorded) if not finite
__device__ __forceinline__ bool cg_alpha(KrylovState* st, const double* ps, int nbs, int it, double& alpha) {
    double v[1];
    sum_partials_all(ps, nbs, v);
    // the reference adds err_tol^2 = 1e-12 to the denominator (poisson2d.cpp:373)
    alpha = (st->mode == ASD2_CG_REFERENCE) ? st->rz[it & 1] / (v[0] + 1.0e-6 * 1.0e-6) : st->rz[it & 1] / v[0];
    const bool ok = isfinite(alpha);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->pq = v[0];
        st->alpha = alpha;
        if (!ok) {
            st->done = DONE_BREAKDOWN;
            st->iter = it;
        }
    }
    return ok;
}

// p_new = z + beta p_old on a tile with a one-cell halo, q = A p_new, partial p_new . q.
// The two direction buffers alternate with the device iteration count (iteration k reads
// P[k & 1] and writes P[(k + 1) & 1]), so iterations skipped while a solve is paused do
// not shift the alternation.
template <typename T, bool MIXED>
__global__ __launch_bounds__(KB) void k_cg_spmv(Geom G, const T* __restrict__ ax, const T* __restrict__ ay,
                                                const T* __restrict__ dg, const T* __restrict__ z, T* p0, T* p1,
                                                T* __restrict__ q, KrylovState* st, const double* pu, int nbu,
                                                double* ps, int it) {
    if (st->done != DONE_RUNNING) return;
    double beta64;
    if (!cg_begin<MIXED>(st, pu, nbu, it, beta64)) return;
    const T beta = static_cast<T>(beta64);
    const bool odd = (it & 1) != 0;
    const T* __restrict__ pold = odd ? p1 : p0;
    T* __restrict__ pnew = odd ? p0 : p1;
    __shared__ T sP[TI + 2][TJ + 2];
    const int ipad = G.nx + 2 * G.g, jpad = G.ny + 2 * G.g;
    const int tiles_i = ceil_div(G.nx, TI), ntiles = tiles_i * ceil_div(G.ny, TJ);
    int t0, t1;
    tile_range(ntiles, t0, t1);
    const int tj = threadIdx.x & (TJ - 1), ti = threadIdx.x / TJ;
    double acc = 0.0;
    for (int t = t0; t < t1; t++) {
        const int bj = t / tiles_i, bi = t - bj * tiles_i;
        const int j0 = G.g + bj * TJ, i0 = G.g + bi * TI;
        for (int e = threadIdx.x; e < (TI + 2) * (TJ + 2); e += KB) {
            const int r = e / (TJ + 2), c = e - r * (TJ + 2);
            const int i = i0 - 1 + r, j = j0 - 1 + c;
            T val = T(0);
            if (i < ipad && j < jpad) {
                const int64_t k = G.idx(i, j);
                val = z[k] + beta * pold[k];
            }
            sP[r][c] = val;
        }
        __syncthreads();
        for (int r = ti; r < TI; r += KB / TJ) {
            const int i = i0 + r, j = j0 + tj;
            if (i >= G.nx + G.g || j >= G.ny + G.g) continue;
            const int64_t k = G.idx(i, j);
            const T pc = sP[r + 1][tj + 1];
            const T qv = dg[k] * pc + ax[k] * sP[r][tj + 1] + ax[k + G.sy] * sP[r + 2][tj + 1] +
                         ay[k] * sP[r + 1][tj] + ay[k + 1] * sP[r + 1][tj + 2];
            pnew[k] = pc;
            q[k] = qv;
            acc += static_cast<double>(pc) * static_cast<double>(qv);
        }
        __syncthreads();
    }
    double v[1] = {acc};
    put_partials(v, ps);
}

// x += alpha p, r -= alpha q, z = M^-1 r, partials r.z and r.r (poisson2d.cpp:346-395).
// p is the direction written by k_cg_spmv of iteration `it` (P[(it + 1) & 1]); pass
// p0 == p1 for a single buffer.
template <typename T>
__global__ __launch_bounds__(KB) void k_cg_update(int64_t n, T* __restrict__ x, const T* p0, const T* p1,
                                                  T* __restrict__ r, const T* __restrict__ q,
                                                  const T* __restrict__ dinv, T* __restrict__ z, KrylovState* st,
                                                  const double* ps, int nbs, double* pu, int it) {
    if (st->done != DONE_RUNNING) return;
    double alpha;
    if (!cg_alpha(st, ps, nbs, it, alpha)) return;
    const T a = static_cast<T>(alpha);
    const T* __restrict__ p = (it & 1) ? p0 : p1;
    double rz = 0.0, rr = 0.0;
    FLAT_LOOP(k, n) {
        const T xv = a * p[k] + x[k];
        const T rv = -a * q[k] + r[k];
        const T zv = dinv[k] * rv;
        x[k] = xv;
        r[k] = rv;
        z[k] = zv;
        rz += static_cast<double>(rv) * static_cast<double>(zv);
        rr += static_cast<double>(rv) * static_cast<double>(rv);
    }
    double v[2] = {rz, rr};
    put_partials(v, pu);
    if (blockIdx.x == 0 && threadIdx.x == 0) st->iter = it + 1;
}

// ------------------------------------------------------------------ mixed-precision reliable updates
// x64 += x32, x32 = 0 (when paused or forced)
__global__ __launch_bounds__(KB) void k_mx_accum(int64_t n, double* __restrict__ x64, float* __restrict__ x32,
                                                 const KrylovState* st, int force) {
    if (!force && st->done != DONE_PAUSED) return;
    FLAT_LOOP(k, n) {
        x64[k] += static_cast<double>(x32[k]);
        x32[k] = 0.0f;
    }
}

// r64 = b - A x64 in FP64 (interior), partial ||r64||^2.  mode 0: reliable update (runs
// while paused), mode 1: report only.
__global__ __launch_bounds__(KB) void k_true_residual(Geom G, const double* __restrict__ ax,
                                                      const double* __restrict__ ay, const double* __restrict__ dg,
                                                      const double* __restrict__ b, const double* __restrict__ x,
                                                      double* __restrict__ r64, const KrylovState* st,
                                                      double* partials, int mode) {
    if (mode == 0 && st->done != DONE_PAUSED) return;
    double acc = 0.0;
    for_interior(G, [&](int, int, int64_t k) {
        const double ax_ = dg[k] * x[k] + ax[k] * x[k - G.sy] + ax[k + G.sy] * x[k + G.sy] + ay[k] * x[k - 1] +
                           ay[k + 1] * x[k + 1];
        const double rv = b[k] - ax_;
        if (r64) r64[k] = rv;
        acc += rv * rv;
    });
    double v[1] = {acc};
    put_partials(v, partials);
}

// r32 = r64, z32 = M^-1 r32, partials r.z and r.r in the layout of k_cg_update (the next
// k_cg_spmv derives beta = rho_true / rho_prev from them, keeping the direction)
__global__ __launch_bounds__(KB) void k_mx_resume(int64_t n, const double* __restrict__ r64, float* __restrict__ r32,
                                                  const float* __restrict__ dinv, float* __restrict__ z32,
                                                  const KrylovState* st, double* partials) {
    if (st->done != DONE_PAUSED) return;
    double rz = 0.0, rr = 0.0;
    FLAT_LOOP(k, n) {
        const float rv = static_cast<float>(r64[k]);
        const float zv = dinv[k] * rv;
        r32[k] = rv;
        z32[k] = zv;
        rz += static_cast<double>(rv) * static_cast<double>(zv);
        rr += static_cast<double>(rv) * static_cast<double>(rv);
    }
    double v[2] = {rz, rr};
    put_partials(v, partials);
}

// ------------------------------------------------------------------ generic pieces (BiCGStab, CSR)
__global__ __launch_bounds__(KB) void k_stencil_apply(Geom G, const double* __restrict__ ax,
                                                      const double* __restrict__ ay, const double* __restrict__ dg,
                                                      const double* __restrict__ x, double* __restrict__ y) {
    for_interior(G, [&](int, int, int64_t k) {
        y[k] = dg[k] * x[k] + ax[k] * x[k - G.sy] + ax[k + G.sy] * x[k + G.sy] + ay[k] * x[k - 1] +
               ay[k + 1] * x[k + 1];
    });
}

// y = A x (stencil) with partials (w . y, y . y); phase 2 is skipped after the half-step exit
__global__ __launch_bounds__(KB) void k_stencil_dot(Geom G, const double* __restrict__ ax,
                                                    const double* __restrict__ ay, const double* __restrict__ dg,
                                                    const double* __restrict__ x, double* __restrict__ y,
                                                    const double* __restrict__ w, const KrylovState* st,
                                                    double* partials, int phase) {
    if (st->done != DONE_RUNNING) return;
    if (phase == 2 && (st->flags & 1)) return;
    double a0 = 0.0, a1 = 0.0;
    for_interior(G, [&](int, int, int64_t k) {
        const double yv = dg[k] * x[k] + ax[k] * x[k - G.sy] + ax[k + G.sy] * x[k + G.sy] + ay[k] * x[k - 1] +
                          ay[k + 1] * x[k + 1];
        y[k] = yv;
        a0 += w[k] * yv;
        a1 += yv * yv;
    });
    double v[2] = {a0, a1};
    put_partials(v, partials);
}

// BiCGStab vector phases (van der Vorst 1992, IML++ ordering)
__global__ __launch_bounds__(KB) void k_bicg_dir(int64_t n, const double* __restrict__ r, double* __restrict__ p,
                                                 const double* __restrict__ v, const double* __restrict__ dinv,
                                                 double* __restrict__ ph, const KrylovState* st) {
    if (st->done != DONE_RUNNING) return;
    const bool first = (st->iter == 0);
    const double beta = st->beta, om = st->omega;
    FLAT_LOOP(k, n) {
        const double pv = first ? r[k] : r[k] + beta * (p[k] - om * v[k]);
        p[k] = pv;
        ph[k] = dinv[k] * pv;
    }
}

__global__ __launch_bounds__(KB) void k_bicg_s(int64_t n, const double* __restrict__ r, const double* __restrict__ v,
                                               const double* __restrict__ dinv, double* __restrict__ s,
                                               double* __restrict__ sh, const KrylovState* st, double* partials) {
    if (st->done != DONE_RUNNING) return;
    const double a = st->alpha;
    double ss = 0.0;
    FLAT_LOOP(k, n) {
        const double sv = r[k] - a * v[k];
        s[k] = sv;
        sh[k] = dinv[k] * sv;
        ss += sv * sv;
    }
    double vv[1] = {ss};
    put_partials(vv, partials);
}

__global__ __launch_bounds__(KB) void k_bicg_x(int64_t n, double* __restrict__ x, const double* __restrict__ ph,
                                               const double* __restrict__ sh, double* __restrict__ r,
                                               const double* __restrict__ s, const double* __restrict__ t,
                                               const double* __restrict__ rt, const KrylovState* st,
                                               double* partials) {
    if (st->done != DONE_RUNNING) return;
    const bool half = (st->flags & 1) != 0;
    const double a = st->alpha, om = half ? 0.0 : st->omega;
    double rr = 0.0, rtr = 0.0;
    FLAT_LOOP(k, n) {
        const double xv = x[k] + a * ph[k] + om * sh[k];
        const double rv = half ? s[k] : s[k] - om * t[k];
        x[k] = xv;
        r[k] = rv;
        rr += rv * rv;
        rtr += rt[k] * rv;
    }
    double v[2] = {rr, rtr};
    put_partials(v, partials);
}

__global__ __launch_bounds__(KB) void k_bicg_init(int64_t n, const double* __restrict__ b, double* __restrict__ x,
                                                  double* __restrict__ r, double* __restrict__ rt, double* partials) {
    double bb = 0.0;
    FLAT_LOOP(k, n) {
        const double bv = b[k];
        x[k] = 0.0;
        r[k] = bv;
        rt[k] = bv;
        bb += bv * bv;
    }
    double v[1] = {bb};
    put_partials(v, partials);
}

// CSR pieces (caller-built matrices, PoissonSolver2D::CG_2FUniform_2D): the recurrence
// step and p = z + beta p in place, then q = A p with partial p.q
__global__ __launch_bounds__(KB) void k_csr_dir(int64_t n, const double* __restrict__ z, double* __restrict__ p,
                                                KrylovState* st, const double* pu, int nbu, int it) {
    if (st->done != DONE_RUNNING) return;
    double beta;
    if (!cg_begin<false>(st, pu, nbu, it, beta)) return;
    FLAT_LOOP(k, n) p[k] = z[k] + beta * p[k];
}

__global__ __launch_bounds__(KB) void k_csr_spmv(int n, const int* __restrict__ rp, const int* __restrict__ cols,
                                                 const double* __restrict__ vals, const double* __restrict__ p,
                                                 double* __restrict__ q, const KrylovState* st, double* partials) {
    if (st->done != DONE_RUNNING) return;
    double acc = 0.0;
    FLAT_LOOP(row, n) {
        double s = 0.0;
        for (int e = rp[row]; e < rp[row + 1]; e++) s += vals[e] * p[cols[e]];
        q[row] = s;
        acc += p[row] * s;
    }
    double v[1] = {acc};
    put_partials(v, partials);
}

// y = A x with partials (w . y, y . y) for BiCGStab (phase 1: alpha, phase 2: omega)
__global__ __launch_bounds__(KB) void k_csr_dot(int n, const int* __restrict__ rp, const int* __restrict__ cols,
                                                const double* __restrict__ vals, const double* __restrict__ x,
                                                double* __restrict__ y, const double* __restrict__ w,
                                                const KrylovState* st, double* partials, int phase) {
    if (st->done != DONE_RUNNING) return;
    if (phase == 2 && (st->flags & 1)) return;
    double a0 = 0.0, a1 = 0.0;
    FLAT_LOOP(row, n) {
        double s = 0.0;
        for (int e = rp[row]; e < rp[row + 1]; e++) s += vals[e] * x[cols[e]];
        y[row] = s;
        a0 += w[row] * s;
        a1 += s * s;
    }
    double v[2] = {a0, a1};
    put_partials(v, partials);
}

__global__ void k_recip(int64_t n, const double* __restrict__ d, double* __restrict__ di) {
    FLAT_LOOP(k, n) di[k] = 1.0 / d[k];
}

// copy the interior of x into the interior of ps
__global__ __launch_bounds__(KB) void k_copy_interior(Geom G, const double* __restrict__ x, double* __restrict__ ps) {
    for_interior(G, [&](int, int, int64_t k) { ps[k] = x[k]; });
}

// ------------------------------------------------------------------ host driver
struct Driver {
    asd2_ctx* ctx;
    KrylovState* st = nullptr;       // device
    KrylovState* host = nullptr;     // pinned, 2 slots
    hipEvent_t ev[2] = {nullptr, nullptr};
    double* pu = nullptr;            // partials of the residual producers (r.z, r.r[, b.b]); tail kernels
    double* ps = nullptr;            // partials of the SpMV producers (p.q [, ...])
};

int flat_blocks(int64_t n) { return std::max(1, std::min<int>(MAX_BLOCKS, ceil_div(n, KB))); }
int spmv_blocks(const Geom& G) {
    return std::max(1, std::min<int>(MAX_BLOCKS, ceil_div(G.nx, TI) * ceil_div(G.ny, TJ)));
}
int interior_blocks(const Geom& G) {
    return std::max(1, std::min<int>(MAX_BLOCKS, ceil_div(G.nx, 4) * ceil_div(G.ny, 64)));
}

template <int OP>
int tail(Driver* D, int nb, const double* partials) {
    hipLaunchKernelGGL(k_tail<OP>, dim3(1), dim3(KB), 0, D->ctx->stream, nb, partials, D->st);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int driver_init(asd2_ctx* ctx, Driver* D) {
    D->ctx = ctx;
    static_assert(sizeof(KrylovState) % 8 == 0, "state size");
    ASD2_TRY(ws(ctx, "kr_state", sizeof(KrylovState) / 8, reinterpret_cast<double**>(&D->st)));
    double* part = nullptr;
    ASD2_TRY(ws(ctx, "kr_partials", 5 * MAX_BLOCKS, &part));
    D->pu = part;
    D->ps = part + 3 * MAX_BLOCKS;
    if (!ctx->kr_host) ASD2_HIP(hipHostMalloc(&ctx->kr_host, 2 * sizeof(KrylovState), hipHostMallocDefault));
    for (int k = 0; k < 2; k++)
        if (!ctx->kr_ev[k]) ASD2_HIP(hipEventCreateWithFlags(&ctx->kr_ev[k], hipEventDisableTiming));
    D->host = static_cast<KrylovState*>(ctx->kr_host);
    D->ev[0] = ctx->kr_ev[0];
    D->ev[1] = ctx->kr_ev[1];
    return ASD2_OK;
}

int set_state(Driver* D, const asd2_solve_opts* o) {
    ASD2_HIP(hipStreamSynchronize(D->ctx->stream));  // the pinned slot may still be a copy target
    KrylovState& s = D->host[0];
    s = KrylovState{};
    s.tol = o->tol;
    s.maxit = o->max_iter > 0 ? o->max_iter : 1000000;
    s.mode = o->mode;
    s.delta = o->reliable_delta > 0 ? o->reliable_delta : 0.1;
    ASD2_HIP(hipMemcpyAsync(D->st, &s, sizeof(s), hipMemcpyHostToDevice, D->ctx->stream));
    return ASD2_OK;
}

int read_state(Driver* D, KrylovState* out) {
    ASD2_HIP(hipMemcpyAsync(&D->host[0], D->st, sizeof(KrylovState), hipMemcpyDeviceToHost, D->ctx->stream));
    ASD2_HIP(hipStreamSynchronize(D->ctx->stream));
    *out = D->host[0];
    return ASD2_OK;
}

// Runs `batch(count)` repeatedly, polling one batch behind, until done is neither
// RUNNING nor (when has_pause) PAUSED after on_pause(state) ran.
template <class Batch, class Pause>
int drive(Driver* D, int check_every, Batch batch, Pause on_pause, bool has_pause, KrylovState* final_state) {
    asd2_ctx* ctx = D->ctx;
    int slot = 0;
    auto enqueue = [&](int s) -> int {
        ASD2_TRY(batch(check_every));
        ASD2_HIP(hipMemcpyAsync(&D->host[s], D->st, sizeof(KrylovState), hipMemcpyDeviceToHost, ctx->stream));
        ASD2_HIP(hipEventRecord(D->ev[s], ctx->stream));
        return ASD2_OK;
    };
    ASD2_TRY(enqueue(0));
    ASD2_TRY(enqueue(1));
    for (int guard = 0; guard < 100000000; guard++) {
        ASD2_HIP(hipEventSynchronize(D->ev[slot]));
        const KrylovState s = D->host[slot];
        if (s.done == DONE_PAUSED && has_pause) {
            ASD2_TRY(on_pause(s));
        } else if (s.done != DONE_RUNNING) {
            break;
        }
        ASD2_TRY(enqueue(slot));
        slot ^= 1;
    }
    return read_state(D, final_state);
}

void fill_stats(const KrylovState& s, asd2_solve_stats* stats) {
    if (!stats) return;
    stats->iterations = s.iter;
    stats->reliable_updates = s.updates;
    stats->fallbacks = s.fallbacks;
    stats->converged = (s.done == DONE_CONVERGED) ? 1 : 0;
    stats->rel_residual = sqrt(s.rr) / s.bnorm;
    stats->true_rel_residual = s.aux2;
}

int status_of(const KrylovState& s) {
    if (s.done == DONE_CONVERGED) return ASD2_OK;
    if (s.done == DONE_BREAKDOWN) {
        set_error("Krylov breakdown (non-finite scalar) after %d iterations", s.iter);
        return ASD2_ERR_NONFINITE;
    }
    set_error("no convergence after %d iterations (||r||/||b|| = %g)", s.iter, sqrt(s.rr) / s.bnorm);
    return ASD2_ERR_NOT_CONVERGED;
}

int report_true_residual(asd2_ctx* ctx, Driver* D, const System& S, const double* x, KrylovState* fs) {
    const Geom& G = ctx->geo;
    const int nb = interior_blocks(G);
    hipLaunchKernelGGL(k_true_residual, dim3(nb), dim3(KB), 0, ctx->stream, G, S.ax, S.ay, S.diag, S.rhs, x,
                       static_cast<double*>(nullptr), D->st, D->ps, 1);
    ASD2_LAUNCH_CHECK();
    ASD2_TRY(tail<T_TRUE_RES_REPORT>(D, nb, D->ps));
    return read_state(D, fs);
}

int solve_fp64(asd2_ctx* ctx, const asd2_solve_opts* o, double* d_ps, asd2_solve_stats* stats) {
    const Geom& G = ctx->geo;
    const int64_t n = G.size();
    System S;
    ASD2_TRY(system_arrays(ctx, &S, false));
    double *x, *r, *z, *p0, *p1, *q;
    ASD2_TRY(ws(ctx, "cg_x", n, &x));
    ASD2_TRY(ws(ctx, "cg_r", n, &r));
    ASD2_TRY(ws(ctx, "cg_z", n, &z));
    ASD2_TRY(ws(ctx, "cg_p0", n, &p0));
    ASD2_TRY(ws(ctx, "cg_p1", n, &p1));
    ASD2_TRY(ws(ctx, "cg_q", n, &q));
    const int nbs = spmv_blocks(G), nbf = flat_blocks(n);
    Driver D;
    ASD2_TRY(driver_init(ctx, &D));
    ASD2_TRY(set_state(&D, o));
    hipLaunchKernelGGL(k_cg_init<double>, dim3(nbf), dim3(KB), 0, ctx->stream, n, S.rhs, x, r,
                       static_cast<double*>(nullptr), static_cast<double*>(nullptr), S.dinv, z, p0, D.pu);
    ASD2_LAUNCH_CHECK();
    ASD2_TRY(tail<T_CG_INIT>(&D, nbf, D.pu));
    int it = 0;
    auto batch = [&](int count) -> int {
        for (int k = 0; k < count; k++, it++) {
            prof_begin(ctx, ASD2_KID_PCG64_SPMV);
            hipLaunchKernelGGL((k_cg_spmv<double, false>), dim3(nbs), dim3(KB), 0, ctx->stream, G, S.ax, S.ay,
                               S.diag, z, p0, p1, q, D.st, D.pu, nbf, D.ps, it);
            prof_end(ctx, ASD2_KID_PCG64_SPMV);
            prof_begin(ctx, ASD2_KID_PCG64_UPDATE);
            hipLaunchKernelGGL(k_cg_update<double>, dim3(nbf), dim3(KB), 0, ctx->stream, n, x, p0, p1, r, q, S.dinv,
                               z, D.st, D.ps, nbs, D.pu, it);
            prof_end(ctx, ASD2_KID_PCG64_UPDATE);
        }
        ASD2_LAUNCH_CHECK();
        return ASD2_OK;
    };
    KrylovState fs;
    ASD2_TRY(drive(&D, o->check_every > 0 ? o->check_every : 32, batch,
                   [](const KrylovState&) { return ASD2_OK; }, false, &fs));
    hipLaunchKernelGGL(k_copy_interior, dim3(interior_blocks(G)), dim3(KB), 0, ctx->stream, G, x, d_ps);
    ASD2_LAUNCH_CHECK();
    const int st = status_of(fs);
    ASD2_TRY(report_true_residual(ctx, &D, S, x, &fs));
    fill_stats(fs, stats);
    return st;
}

int solve_mixed(asd2_ctx* ctx, const asd2_solve_opts* o, double* d_ps, asd2_solve_stats* stats) {
    const Geom& G = ctx->geo;
    const int64_t n = G.size();
    System S;
    ASD2_TRY(system_arrays(ctx, &S, false));
    ASD2_TRY(ensure_system32(ctx, &S));
    double *x64, *r64;
    float *x, *r, *z, *p0, *p1, *q;
    ASD2_TRY(ws(ctx, "cg_x", n, &x64));
    ASD2_TRY(ws(ctx, "cg_r", n, &r64));
    ASD2_TRY(ws(ctx, "mx_x", n, &x));
    ASD2_TRY(ws(ctx, "mx_r", n, &r));
    ASD2_TRY(ws(ctx, "mx_z", n, &z));
    ASD2_TRY(ws(ctx, "mx_p0", n, &p0));
    ASD2_TRY(ws(ctx, "mx_p1", n, &p1));
    ASD2_TRY(ws(ctx, "mx_q", n, &q));
    const int nbs = spmv_blocks(G), nbf = flat_blocks(n), nbi = interior_blocks(G);
    Driver D;
    ASD2_TRY(driver_init(ctx, &D));
    ASD2_TRY(set_state(&D, o));
    hipLaunchKernelGGL(k_cg_init<float>, dim3(nbf), dim3(KB), 0, ctx->stream, n, S.rhs, x, r, r64, x64, S.dinv32, z,
                       p0, D.pu);
    ASD2_LAUNCH_CHECK();
    ASD2_TRY(tail<T_CG_INIT>(&D, nbf, D.pu));
    int it = 0;
    auto batch = [&](int count) -> int {
        for (int k = 0; k < count; k++, it++) {
            prof_begin(ctx, ASD2_KID_PCG32_SPMV);
            hipLaunchKernelGGL((k_cg_spmv<float, true>), dim3(nbs), dim3(KB), 0, ctx->stream, G, S.ax32, S.ay32,
                               S.diag32, z, p0, p1, q, D.st, D.pu, nbf, D.ps, it);
            prof_end(ctx, ASD2_KID_PCG32_SPMV);
            prof_begin(ctx, ASD2_KID_PCG32_UPDATE);
            hipLaunchKernelGGL(k_cg_update<float>, dim3(nbf), dim3(KB), 0, ctx->stream, n, x, p0, p1, r, q,
                               S.dinv32, z, D.st, D.ps, nbs, D.pu, it);
            prof_end(ctx, ASD2_KID_PCG32_UPDATE);
        }
        ASD2_LAUNCH_CHECK();
        return ASD2_OK;
    };
    // Reliable update at the paused iteration k: fold x32 into x64, r64 = b - A x64, replace
    // the FP32 residual, and re-run iteration k (its SpMV derives beta from the new r.z).
    auto pause = [&](const KrylovState& s) -> int {
        it = s.iter;
        hipLaunchKernelGGL(k_mx_accum, dim3(nbf), dim3(KB), 0, ctx->stream, n, x64, x, D.st, 0);
        hipLaunchKernelGGL(k_true_residual, dim3(nbi), dim3(KB), 0, ctx->stream, G, S.ax, S.ay, S.diag, S.rhs, x64,
                           r64, D.st, D.ps, 0);
        ASD2_TRY(tail<T_TRUE_RES>(&D, nbi, D.ps));
        hipLaunchKernelGGL(k_mx_resume, dim3(nbf), dim3(KB), 0, ctx->stream, n, r64, r, S.dinv32, z, D.st, D.pu);
        ASD2_TRY(tail<T_MX_RESUME>(&D, 0, D.pu));
        ASD2_LAUNCH_CHECK();
        return ASD2_OK;
    };
    KrylovState fs;
    ASD2_TRY(drive(&D, o->check_every > 0 ? o->check_every : 16, batch, pause, true, &fs));
    // fold any FP32 progress not yet accumulated (max-iteration exit)
    hipLaunchKernelGGL(k_mx_accum, dim3(nbf), dim3(KB), 0, ctx->stream, n, x64, x, D.st, 1);
    hipLaunchKernelGGL(k_copy_interior, dim3(nbi), dim3(KB), 0, ctx->stream, G, x64, d_ps);
    ASD2_LAUNCH_CHECK();
    const int st = status_of(fs);
    ASD2_TRY(report_true_residual(ctx, &D, S, x64, &fs));
    fill_stats(fs, stats);
    return st;
}

}  // namespace asd2

using namespace asd2;

extern "C" {

int asd2_gfm_apply(asd2_ctx* ctx, const double* d_x, double* d_y) {
    ASD2_TRY(check_ctx(ctx));
    if (!ctx->system_ready) { set_error("no assembled system"); return ASD2_ERR_STATE; }
    if (!d_x || !d_y) { set_error("null array"); return ASD2_ERR_ARG; }
    System S;
    ASD2_TRY(system_arrays(ctx, &S, false));
    const Geom& G = ctx->geo;
    hipLaunchKernelGGL(k_stencil_apply, dim3(interior_blocks(G)), dim3(KB), 0, ctx->stream, G, S.ax, S.ay, S.diag,
                       d_x, d_y);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_poisson_solve(asd2_ctx* ctx, const asd2_solve_opts* opts, double* d_ps, asd2_solve_stats* stats) {
    ASD2_TRY(check_ctx(ctx));
    if (!opts || !d_ps) { set_error("null argument"); return ASD2_ERR_ARG; }
    if (!ctx->system_ready) { set_error("no assembled system (call asd2_gfm_assemble first)"); return ASD2_ERR_STATE; }
    if (opts->mode != ASD2_CG_REFERENCE && opts->mode != ASD2_CG_STANDARD) {
        set_error("bad CG mode %d", opts->mode);
        return ASD2_ERR_ARG;
    }
    if (opts->precond == ASD2_PRECOND_LEARNED) return learned_solve(ctx, opts, d_ps, stats);
    if (opts->precision == ASD2_PREC_MIXED) return solve_mixed(ctx, opts, d_ps, stats);
    return solve_fp64(ctx, opts, d_ps, stats);
}

int asd2_poisson_bicgstab(asd2_ctx* ctx, const asd2_solve_opts* o, double* d_ps, asd2_solve_stats* stats) {
    ASD2_TRY(check_ctx(ctx));
    if (!o || !d_ps) { set_error("null argument"); return ASD2_ERR_ARG; }
    if (!ctx->system_ready) { set_error("no assembled system"); return ASD2_ERR_STATE; }
    const Geom& G = ctx->geo;
    const int64_t n = G.size();
    System S;
    ASD2_TRY(system_arrays(ctx, &S, false));
    double *x, *r, *rt, *p, *v, *ph, *s, *sh, *t;
    ASD2_TRY(ws(ctx, "cg_x", n, &x));
    ASD2_TRY(ws(ctx, "cg_r", n, &r));
    ASD2_TRY(ws(ctx, "bi_rt", n, &rt));
    ASD2_TRY(ws(ctx, "bi_p", n, &p));
    ASD2_TRY(ws(ctx, "bi_v", n, &v));
    ASD2_TRY(ws(ctx, "bi_ph", n, &ph));
    ASD2_TRY(ws(ctx, "bi_s", n, &s));
    ASD2_TRY(ws(ctx, "bi_sh", n, &sh));
    ASD2_TRY(ws(ctx, "bi_t", n, &t));
    const int nbf = flat_blocks(n), nbi = interior_blocks(G);
    Driver D;
    ASD2_TRY(driver_init(ctx, &D));
    ASD2_TRY(set_state(&D, o));
    hipLaunchKernelGGL(k_bicg_init, dim3(nbf), dim3(KB), 0, ctx->stream, n, S.rhs, x, r, rt, D.pu);
    ASD2_LAUNCH_CHECK();
    ASD2_TRY(tail<T_BICG_INIT>(&D, nbf, D.pu));
    auto batch = [&](int count) -> int {
        for (int k = 0; k < count; k++) {
            hipLaunchKernelGGL(k_bicg_dir, dim3(nbf), dim3(KB), 0, ctx->stream, n, r, p, v, S.dinv, ph, D.st);
            hipLaunchKernelGGL(k_stencil_dot, dim3(nbi), dim3(KB), 0, ctx->stream, G, S.ax, S.ay, S.diag, ph, v, rt,
                               D.st, D.pu, 1);
            ASD2_TRY(tail<T_BICG_ALPHA>(&D, nbi, D.pu));
            hipLaunchKernelGGL(k_bicg_s, dim3(nbf), dim3(KB), 0, ctx->stream, n, r, v, S.dinv, s, sh, D.st,
                               D.pu);
            ASD2_TRY(tail<T_BICG_S>(&D, nbf, D.pu));
            hipLaunchKernelGGL(k_stencil_dot, dim3(nbi), dim3(KB), 0, ctx->stream, G, S.ax, S.ay, S.diag, sh, t, s,
                               D.st, D.pu, 2);
            ASD2_TRY(tail<T_BICG_OMEGA>(&D, nbi, D.pu));
            hipLaunchKernelGGL(k_bicg_x, dim3(nbf), dim3(KB), 0, ctx->stream, n, x, ph, sh, r, s, t, rt, D.st,
                               D.pu);
            ASD2_TRY(tail<T_BICG_X>(&D, nbf, D.pu));
        }
        ASD2_LAUNCH_CHECK();
        return ASD2_OK;
    };
    KrylovState fs;
    ASD2_TRY(drive(&D, o->check_every > 0 ? o->check_every : 16, batch,
                   [](const KrylovState&) { return ASD2_OK; }, false, &fs));
    hipLaunchKernelGGL(k_copy_interior, dim3(nbi), dim3(KB), 0, ctx->stream, G, x, d_ps);
    ASD2_LAUNCH_CHECK();
    const int st = status_of(fs);
    ASD2_TRY(report_true_residual(ctx, &D, S, x, &fs));
    fill_stats(fs, stats);
    return st;
}

int asd2_csr_pcg(asd2_ctx* ctx, int n, const int* d_rowptr, const int* d_cols, const double* d_vals,
                 const double* d_mdiag, const double* d_b, double* d_x, const asd2_solve_opts* o,
                 asd2_solve_stats* stats) {
    ASD2_TRY(check_ctx(ctx));
    if (n <= 0 || !d_rowptr || !d_cols || !d_vals || !d_mdiag || !d_b || !d_x || !o) {
        set_error("bad CSR argument");
        return ASD2_ERR_ARG;
    }
    if (o->mode != ASD2_CG_REFERENCE && o->mode != ASD2_CG_STANDARD) {
        set_error("bad CG mode %d", o->mode);
        return ASD2_ERR_ARG;
    }
    double *r, *z, *p, *q, *dinv;
    ASD2_TRY(ws(ctx, "csr_r", n, &r));
    ASD2_TRY(ws(ctx, "csr_z", n, &z));
    ASD2_TRY(ws(ctx, "csr_p", n, &p));
    ASD2_TRY(ws(ctx, "csr_q", n, &q));
    ASD2_TRY(ws(ctx, "csr_dinv", n, &dinv));
    const int nb = flat_blocks(n);
    const int64_t n64 = n;
    Driver D;
    ASD2_TRY(driver_init(ctx, &D));
    ASD2_TRY(set_state(&D, o));
    hipLaunchKernelGGL(k_recip, dim3(nb), dim3(KB), 0, ctx->stream, n64, d_mdiag, dinv);
    hipLaunchKernelGGL(k_cg_init<double>, dim3(nb), dim3(KB), 0, ctx->stream, n64, d_b, d_x, r,
                       static_cast<double*>(nullptr), static_cast<double*>(nullptr), dinv, z, p, D.pu);
    ASD2_LAUNCH_CHECK();
    ASD2_TRY(tail<T_CG_INIT>(&D, nb, D.pu));
    int it = 0;
    auto batch = [&](int count) -> int {
        for (int k = 0; k < count; k++, it++) {
            hipLaunchKernelGGL(k_csr_dir, dim3(nb), dim3(KB), 0, ctx->stream, n64, z, p, D.st, D.pu, nb, it);
            hipLaunchKernelGGL(k_csr_spmv, dim3(nb), dim3(KB), 0, ctx->stream, n, d_rowptr, d_cols, d_vals, p, q,
                               D.st, D.ps);
            hipLaunchKernelGGL(k_cg_update<double>, dim3(nb), dim3(KB), 0, ctx->stream, n64, d_x, p, p, r, q, dinv,
                               z, D.st, D.ps, nb, D.pu, it);
        }
        ASD2_LAUNCH_CHECK();
        return ASD2_OK;
    };
    KrylovState fs;
    ASD2_TRY(drive(&D, o->check_every > 0 ? o->check_every : 32, batch,
                   [](const KrylovState&) { return ASD2_OK; }, false, &fs));
    fs.aux2 = -1.0;
    fill_stats(fs, stats);
    return status_of(fs);
}

int asd2_csr_bicgstab(asd2_ctx* ctx, int n, const int* d_rowptr, const int* d_cols, const double* d_vals,
                      const double* d_mdiag, const double* d_b, double* d_x, const asd2_solve_opts* o,
                      asd2_solve_stats* stats) {
    ASD2_TRY(check_ctx(ctx));
    if (n <= 0 || !d_rowptr || !d_cols || !d_vals || !d_mdiag || !d_b || !d_x || !o) {
        set_error("bad CSR argument");
        return ASD2_ERR_ARG;
    }
    double *r, *rt, *p, *v, *ph, *s, *sh, *t, *dinv;
    ASD2_TRY(ws(ctx, "csr_r", n, &r));
    ASD2_TRY(ws(ctx, "csr_rt", n, &rt));
    ASD2_TRY(ws(ctx, "csr_p", n, &p));
    ASD2_TRY(ws(ctx, "csr_v", n, &v));
    ASD2_TRY(ws(ctx, "csr_ph", n, &ph));
    ASD2_TRY(ws(ctx, "csr_s", n, &s));
    ASD2_TRY(ws(ctx, "csr_sh", n, &sh));
    ASD2_TRY(ws(ctx, "csr_t", n, &t));
    ASD2_TRY(ws(ctx, "csr_dinv", n, &dinv));
    const int nb = flat_blocks(n);
    const int64_t n64 = n;
    Driver D;
    ASD2_TRY(driver_init(ctx, &D));
    ASD2_TRY(set_state(&D, o));
    hipLaunchKernelGGL(k_recip, dim3(nb), dim3(KB), 0, ctx->stream, n64, d_mdiag, dinv);
    hipLaunchKernelGGL(k_bicg_init, dim3(nb), dim3(KB), 0, ctx->stream, n64, d_b, d_x, r, rt, D.pu);
    ASD2_LAUNCH_CHECK();
    ASD2_TRY(tail<T_BICG_INIT>(&D, nb, D.pu));
    auto batch = [&](int count) -> int {
        for (int k = 0; k < count; k++) {
            hipLaunchKernelGGL(k_bicg_dir, dim3(nb), dim3(KB), 0, ctx->stream, n64, r, p, v, dinv, ph, D.st);
            hipLaunchKernelGGL(k_csr_dot, dim3(nb), dim3(KB), 0, ctx->stream, n, d_rowptr, d_cols, d_vals, ph, v, rt,
                               D.st, D.pu, 1);
            ASD2_TRY(tail<T_BICG_ALPHA>(&D, nb, D.pu));
            hipLaunchKernelGGL(k_bicg_s, dim3(nb), dim3(KB), 0, ctx->stream, n64, r, v, dinv, s, sh, D.st,
                               D.pu);
            ASD2_TRY(tail<T_BICG_S>(&D, nb, D.pu));
            hipLaunchKernelGGL(k_csr_dot, dim3(nb), dim3(KB), 0, ctx->stream, n, d_rowptr, d_cols, d_vals, sh, t, s,
                               D.st, D.pu, 2);
            ASD2_TRY(tail<T_BICG_OMEGA>(&D, nb, D.pu));
            hipLaunchKernelGGL(k_bicg_x, dim3(nb), dim3(KB), 0, ctx->stream, n64, d_x, ph, sh, r, s, t, rt, D.st,
                               D.pu);
            ASD2_TRY(tail<T_BICG_X>(&D, nb, D.pu));
        }
        ASD2_LAUNCH_CHECK();
        return ASD2_OK;
    };
    KrylovState fs;
    ASD2_TRY(drive(&D, o->check_every > 0 ? o->check_every : 16, batch,
                   [](const KrylovState&) { return ASD2_OK; }, false, &fs));
    fs.aux2 = -1.0;
    fill_stats(fs, stats);
    return status_of(fs);
}

}  // extern "C"
