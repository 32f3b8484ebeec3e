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
#include "unet.h"

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

// sum_partials_all of three partial arrays at once (NA values of na partials from pa, then NB
// of nb from pb, then NC of nc from pc; pc == nullptr gives zeros): the loads of all rows are in
// flight together and one block reduction serves them all.  Per value the order is that of
// sum_partials_all.
template <int NA, int NB, int NC>
__device__ __forceinline__ void sum_partials3(const double* pa, int na, const double* pb, int nb, const double* pc,
                                              int nc, double (&v)[NA + NB + NC]) {
    __shared__ double bc[NA + NB + NC];
#pragma unroll
    for (int k = 0; k < NA; k++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < na; b += KB) s += pa[k * na + b];
        v[k] = s;
    }
#pragma unroll
    for (int k = 0; k < NB; k++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += KB) s += pb[k * nb + b];
        v[NA + k] = s;
    }
#pragma unroll
    for (int k = 0; k < NC; k++) {
        double s = 0.0;
        if (pc)
            for (int b = threadIdx.x; b < nc; b += KB) s += pc[k * nc + b];
        v[NA + NB + k] = s;
    }
    block_sum(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NA + NB + NC; k++) bc[k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NA + NB + NC; k++) v[k] = bc[k];
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
    T_FCG_INIT,         // r.r, (D^-1 r)^2, b.b -> bnorm, initial convergence
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
template <> struct TailNV<T_FCG_INIT> { static constexpr int value = 3; };
template <> struct TailNV<T_BICG_OMEGA> { static constexpr int value = 2; };
template <> struct TailNV<T_BICG_X> { static constexpr int value = 2; };

// whether the producer of OP's partials ran (the tail mirrors its guard)
template <int OP>
__device__ __forceinline__ bool tail_active(const KrylovState* st) {
    if (OP == T_CG_INIT || OP == T_FCG_INIT || OP == T_BICG_INIT || OP == T_TRUE_RES_REPORT) return true;
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
    if (OP == T_CG_INIT || OP == T_FCG_INIT) {
        // CG partials are (r.z, r.r, b.b), FCG partials (r.r, (D^-1 r)^2, b.b)
        if (OP == T_FCG_INIT) v[1] = v[0];
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
            T qv;
            if constexpr (sizeof(T) == 4) {  // FP32: flux form, dg = the remainder d (solver.h)
                qv = ax[k] * (sP[r][tj + 1] - pc) + ax[k + G.sy] * (sP[r + 2][tj + 1] - pc) +
                     ay[k] * (sP[r + 1][tj] - pc) + ay[k + 1] * (sP[r + 1][tj + 2] - pc) + dg[k] * pc;
            } else {
                qv = dg[k] * pc + ax[k] * sP[r][tj + 1] + ax[k + G.sy] * sP[r + 2][tj + 1] +
                     ay[k] * sP[r + 1][tj] + ay[k + 1] * sP[r + 1][tj + 2];
            }
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
    double* pz = nullptr;            // partials of the preconditioner output (z.r, z.q)
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
    ASD2_TRY(ws(ctx, "kr_partials", 7 * MAX_BLOCKS, &part));
    D->pu = part;
    D->ps = part + 3 * MAX_BLOCKS;
    D->pz = part + 5 * MAX_BLOCKS;
    if (!ctx->kr_host) ASD2_HIP(hipHostMalloc(&ctx->kr_host, 2 * sizeof(KrylovState), hipHostMallocDefault));
    for (int k = 0; k < 2; k++)
        if (!ctx->kr_ev[k]) ASD2_HIP(hipEventCreateWithFlags(&ctx->kr_ev[k], hipEventDisableTiming));
    D->host = static_cast<KrylovState*>(ctx->kr_host);
    D->ev[0] = ctx->kr_ev[0];
    D->ev[1] = ctx->kr_ev[1];
    return ASD2_OK;
}

// delta0: the reliable-update factor when the options leave it at 0
int set_state(Driver* D, const asd2_solve_opts* o, double delta0 = 0.1) {
    ASD2_HIP(hipStreamSynchronize(D->ctx->stream));  // the pinned slot may still be a copy target
    KrylovState& s = D->host[0];
    s = KrylovState{};
    s.tol = o->tol;
    s.maxit = o->max_iter > 0 ? o->max_iter : 1000000;
    s.mode = o->mode;
    s.delta = o->reliable_delta > 0 ? o->reliable_delta : delta0;
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
// RUNNING nor (when has_pause) PAUSED after on_pause(state) ran.  The batch in the other slot
// was enqueued before on_pause's kernels: it ran as no-ops on the paused state and its
// snapshot is stale, so that snapshot is skipped once.
template <class Batch, class Pause>
int drive(Driver* D, int check_every, Batch batch, Pause on_pause, bool has_pause, KrylovState* final_state) {
    asd2_ctx* ctx = D->ctx;
    int slot = 0;
    bool stale[2] = {false, false};
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
        if (stale[slot]) {
            stale[slot] = false;
        } else if (s.done == DONE_PAUSED && has_pause) {
            ASD2_TRY(on_pause(s));
            stale[slot ^ 1] = true;
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


// ------------------------------------------------------------------ flexible PCG, learned preconditioner
// FP32 flexible CG (Notay 2000): beta_k = (z_k, r_k - r_{k-1}) / (z_{k-1}, r_{k-1})
// = -alpha_{k-1} (z_k, q_{k-1}) / rho_{k-1}, with the preconditioner
//     z = D^-1 r + gamma P y,  y = net(g),  g = P^T (r - A D^-1 r),  gamma = (y.g) / (y.A_c y)
// on c x c blocks (P piecewise constant, A_c = P^T A P), or z = D^-1 r (Jacobi fallback).  The
// network runs in its linear form: the ReLU pattern of a pass on the coarse coefficients, taken
// once per solve, then per iteration a bias-free masked pass on g, so g -> y is one linear map
// for the whole solve and only the scalar gamma varies (a network that is nonlinear in r makes
// every iteration's preconditioner different, and the truncated flexible recurrence then loses
// the Krylov acceleration for every component the network does not capture).  FP64
// reliable updates and the FP64 true residual as in the mixed-precision PCG.  Every scalar of z
// splits into a fine Jacobi part and a coarse part (z.r = r.D^-1 r + gamma y.P^T r, likewise
// z.q_prev), so z is never stored: the SpMV forms it on the fly.  Each iteration: begin
// (decision; one fine pass for g, P^T r, P^T q_prev and the Jacobi dots), pack (network input),
// network, coarse dots (y.g, y.A_c y, y.P^T r, y.P^T q_prev), direction/SpMV, x/r update.

// coarse data in FP64: a coarse diagonal is the sum of c^2 fine diagonals minus twice the
// faces inside the block, a cancellation that loses about log2(c) bits per level, and the
// smooth modes the coarse correction is for sit at the bottom of the spectrum
constexpr int FCG_M = 4;              // directions kept by the truncated flexible recurrence
constexpr int FCG_SLOTS = FCG_M + 1;  // ring of directions: iteration k in slot k % FCG_SLOTS
// directions kept when the network solves a level below the V-cycle's smoothed levels: its
// output varies with the residual through the scalar step on its level, and the network's map is
// not symmetric; four directions took 4x fewer iterations than one on 2048^2 droplet systems
constexpr int HYBRID_M = FCG_M;

struct Coarse {
    int c, Hc, Wc;
    double *acx, *acy, *dc;   // coarse operator, (Hc + 1) x (Wc + 1)
    double *g, *rc;           // P^T (r - A D^-1 r), P^T r, Hc x Wc
    double* qc[FCG_M];        // P^T q of the last FCG_M iterations (j = 1: the latest)
    double* aq;               // P^T A D^-1 q_1 (the multiplicative fine level's coarse dot operand)
};

// directions p_k and q_k = A p_k of the last FCG_SLOTS iterations
struct Ring {
    float* p[FCG_SLOTS];
    float* q[FCG_SLOTS];
};

// the ring seen from iteration k (set on the host, so kernels index it with constants):
// p[j], q[j] = direction j + 1 iterations back, pnew / qnew = the slots of iteration k
struct Dirs {
    const float* p[FCG_M];
    const float* q[FCG_M];
    float* pnew;
    float* qnew;
};

inline Dirs dirs_for(const Ring& R, int k) {
    Dirs d{};
    for (int j = 0; j < FCG_M; j++) {
        const int s = ((k - 1 - j) % FCG_SLOTS + FCG_SLOTS) % FCG_SLOTS;
        d.p[j] = R.p[s];
        d.q[j] = R.q[s];
    }
    d.pnew = R.p[k % FCG_SLOTS];
    d.qnew = R.q[k % FCG_SLOTS];
    return d;
}

// fine region of the coarse fine pass: whole c x c blocks, 32 x 64 cells (c x 64 when c > 32)
__host__ __device__ constexpr int lp_ri(int c) { return c > 32 ? c : 32; }
constexpr int LP_RJ = 64;

__global__ __launch_bounds__(KB) void k_fcg_init(int64_t n, const double* __restrict__ b, float* __restrict__ x,
                                                 float* __restrict__ r, double* __restrict__ r64,
                                                 double* __restrict__ x64, const float* __restrict__ dinv, Ring R,
                                                 double* partials) {
    double rr = 0.0, zz = 0.0, bb = 0.0;
    FLAT_LOOP(k, n) {
        const double bv = b[k];
        const float rv = static_cast<float>(bv);
        const float zv = dinv[k] * rv;
        x[k] = 0.0f;
        r[k] = rv;
        r64[k] = bv;
        x64[k] = 0.0;
#pragma unroll
        for (int s = 0; s < FCG_SLOTS; s++) {
            R.p[s][k] = 0.0f;
            R.q[s][k] = 0.0f;
        }
        rr += static_cast<double>(rv) * rv;
        zz += static_cast<double>(zv) * zv;
        bb += bv * bv;
    }
    double v[3] = {rr, zz, bb};
    put_partials(v, partials);
}

__global__ __launch_bounds__(KB) void k_lp_coarse_op(Geom G, const double* __restrict__ ax,
                                                     const double* __restrict__ ay, const double* __restrict__ dg,
                                                     Coarse C) {
    const int64_t cs = static_cast<int64_t>(C.Hc + 1) * (C.Wc + 1);
    FLAT_LOOP(e, cs) {
        const int I = static_cast<int>(e / (C.Wc + 1)), J = static_cast<int>(e - static_cast<int64_t>(I) * (C.Wc + 1));
        coarse_cell(ax, ay, dg, G.sy, G.g, C.c, C.Hc, C.Wc, I, J, C.acx, C.acy, C.dc);
    }
}

// ReLU-pass input (0, log coarse W-face, S-face coefficients, log coarse diagonal) and the
// linear-pass input cleared (its channel 0 is written per iteration)
__global__ __launch_bounds__(KB) void k_lp_static_features(Coarse C, uint16_t* __restrict__ feat,
                                                           uint16_t* __restrict__ flin, float h2x, float h2y) {
    const int64_t Pc = static_cast<int64_t>(C.Hc) * C.Wc;
    const float h2 = fminf(h2x, h2y);
    FLAT_LOOP(p, Pc) {
        const int I = static_cast<int>(p / C.Wc), J = static_cast<int>(p - static_cast<int64_t>(I) * C.Wc);
        const int k = I * (C.Wc + 1) + J;
        const uint32_t lo = static_cast<uint32_t>(f2bf(log_coeff(static_cast<float>(C.acx[k]), h2x))) << 16;
        const uint32_t hi = static_cast<uint32_t>(f2bf(log_coeff(static_cast<float>(C.acy[k]), h2y))) |
                            (static_cast<uint32_t>(f2bf(log_diag(static_cast<float>(C.dc[k]), h2))) << 16);
        *reinterpret_cast<uint2*>(feat + p * 4) = make_uint2(lo, hi);
        *reinterpret_cast<uint2*>(flin + p * 4) = make_uint2(0u, 0u);
    }
}

// sum of v over the c consecutive lanes of a block row (c a power of two)
__device__ __forceinline__ double seg_sum(double v, int c) {
    for (int m = 1; m < c && m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// alpha = rho / p.q of iteration `it`, p.q recorded in its ring slot
__device__ __forceinline__ bool fcg_alpha(KrylovState* st, const double* ps, int nbs, int it, double& alpha) {
    double v[1];
    sum_partials_all(ps, nbs, v);
    alpha = st->rz[it & 1] / v[0];
    const bool ok = isfinite(alpha);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->pq = v[0];
        st->pqv[it % FCG_SLOTS] = v[0];
        st->alpha = alpha;
        if (!ok) {
            st->done = DONE_BREAKDOWN;
            st->iter = it;
        }
    }
    return ok;
}

// The fine pass of the coarse correction for iteration `itn`, over regions of whole c x c
// blocks: g = P^T (r - A D^-1 r), P^T r and P^T q_j of the last nwin directions into the
// coarse arrays, partials (r.D^-1 r, (D^-1 r).q_1 .. (D^-1 r).q_FCG_M, sum (g / dc)^2).
// The diagonal term of A D^-1 r is r itself, so r - A D^-1 r = -(off-diagonal terms of A applied
// to D^-1 r) and the diagonal is never read.
// UPD = false: r = r_in (the first iteration after a start or a reliable update).  UPD = true:
// fused with the update of iteration itn - 1 (alpha = rho / p.q, x += alpha p, r_out = r_in -
// alpha q, partials (r.r, 0) for the next decision), so the sums are those of the new residual.
// r alternates between two buffers (neighbouring regions read the old r of their halo while
// the owner writes the new one).
// MULT (the multiplicative fine level, one direction): z0 = omega D^-1 r before the coarse
// correction, so g = P^T (r - omega A D^-1 r); with A D^-1 r = r + off, also the rows
// (A D^-1 r).D^-1 r and (A D^-1 r).D^-1 q_1 into pm and the coarse vector P^T A D^-1 q_1 into C.aq
// (the fine parts of the smoothed z's dots, see k_lp_spmv_mult).
template <bool UPD, int CC, int MW, bool MULT = false>
__global__ __launch_bounds__(KB) void k_lp_pass(Geom G, Coarse C, const float* __restrict__ r_in,
                                                float* __restrict__ r_out, float* __restrict__ x, Dirs Dd,
                                                const float* __restrict__ dinv, const float* __restrict__ ax,
                                                const float* __restrict__ ay, KrylovState* st, const double* ps,
                                                int nbs, double* pu, int itn,
                                                int nwin, double* pb, KrylovState* snap, double omega = 0.0,
                                                double* pm = nullptr) {
    static_assert(!MULT || MW == 1, "the multiplicative fine level keeps one direction");
    // snap (pinned host memory, or nullptr): the state as this kernel leaves it, written by the
    // thread that owns the state's updates, for the host's poll after the launch
    auto snapshot = [&]() {
        if (snap && blockIdx.x == 0 && threadIdx.x == 0) *snap = *st;
    };
    if (st->done != DONE_RUNNING) {
        snapshot();
        return;
    }
    // Dd is the ring seen from itn: Dd.p[0], Dd.q[0] = the direction of the updated iteration
    const float* __restrict__ pu_dir = Dd.p[0];
    const float* __restrict__ qu_dir = Dd.q[0];
    const float* qj[MW];
#pragma unroll
    for (int j = 0; j < MW; j++) qj[j] = Dd.q[j];
    constexpr int c = CC, RI = lp_ri(CC), nbi = RI / c, nbj = LP_RJ / c, nblk = nbi * nbj;
    constexpr int NS = 2 + MW + (MULT ? 1 : 0);  // block sums: w, r, q_1 .. q_MW (, A D^-1 q_1)
    // Wave k handles the region rows k, k + NWV, ...  With c >= 4 each lane sums its column over
    // the RPW rows of a block that its wave handles, the segment sum adds the c columns, and the
    // result goes to the wave's own slot (written once); the slots are added in wave order, so
    // the sums do not depend on timing.  With c <= 2 a block gets at most two row sums, and
    // x + y = y + x, so one slot with atomics is exact too.
    constexpr int NWV = KB / 64, NSLOT = CC >= 4 ? NWV : 1, RPW = CC >= 4 ? CC / NWV : 1;
    __shared__ float sZ[RI + 2][LP_RJ + 2];
    __shared__ float sR[RI][LP_RJ];
    __shared__ float sQ[UPD ? RI : 1][LP_RJ];  // q of the updated direction (= q_1) on the region
    __shared__ float sDQ[MULT ? RI + 2 : 1][MULT ? LP_RJ + 2 : 1];  // D^-1 q_1 on the region and halo
    __shared__ double sB[NSLOT][NS][nblk];
    const bool mq = MULT && nwin >= 1;  // the q_1 terms of the multiplicative level
    const int reg_i = ceil_div(G.nx, RI), nreg = reg_i * ceil_div(G.ny, LP_RJ);
    const int ilo = G.g, ihi = G.g + G.nx, jlo = G.g, jhi = G.g + G.ny;
    const int lane = threadIdx.x & 63;
    int t0, t1;
    tile_range(nreg, t0, t1);
    float alpha = 0.f;
    const int it = itn - 1;  // the iteration updated when UPD
    if (UPD) {
        double a64;
        if (!fcg_alpha(st, ps, nbs, it, a64)) {
            snapshot();
            return;
        }
        alpha = static_cast<float>(a64);
    }
    double a_rz = 0.0, a_uu = 0.0, a_rr = 0.0, a_zq[MW] = {}, a_ss = 0.0, a_sq = 0.0;
    for (int t = t0; t < t1; t++) {
        const int rj = t / reg_i, ri = t - rj * reg_i;
        const int i0 = ilo + ri * RI, j0 = jlo + rj * LP_RJ;
        for (int e = threadIdx.x; e < (RI + 2) * (LP_RJ + 2); e += KB) {
            const int a = e / (LP_RJ + 2), b = e - a * (LP_RJ + 2);
            const int i = i0 - 1 + a, j = j0 - 1 + b;
            float v = 0.0f, dq = 0.0f;
            if (i >= ilo && i < ihi && j >= jlo && j < jhi) {
                const int64_t k = G.idx(i, j);
                float rv = r_in[k];
                float qv = 0.0f;
                if (UPD) {
                    qv = qu_dir[k];
                    rv = -alpha * qv + rv;
                } else if (mq) {
                    qv = qj[0][k];
                }
                const float dk = dinv[k];
                v = dk * rv;
                if (MULT) dq = dk * qv;
                if (a >= 1 && a <= RI && b >= 1 && b <= LP_RJ) {
                    sR[a - 1][b - 1] = rv;
                    if (UPD) sQ[UPD ? a - 1 : 0][b - 1] = qv;
                }
            }
            sZ[a][b] = v;
            if (MULT) sDQ[MULT ? a : 0][MULT ? b : 0] = dq;
        }
        if constexpr (NSLOT == 1)
            for (int e = threadIdx.x; e < NS * nblk; e += KB) (&sB[0][0][0])[e] = 0.0;
        __syncthreads();
        double cw = 0.0, cr = 0.0, cq[MW] = {}, caq = 0.0;  // the lane's column sums over its rows of a block
        for (int e = threadIdx.x, it_r = 0; e < RI * LP_RJ; e += KB, it_r++) {
            const int a = e / LP_RJ, b = e - a * LP_RJ;  // a wave covers one region row
            const int i = i0 + a, j = j0 + b;
            const bool in = i < ihi && j < jhi;
            double w = 0.0, rv = 0.0, qv[MW] = {}, aqv = 0.0;  // FP64 sums of FP32 data (the inputs are exact)
            if (in) {
                const int64_t k = G.idx(i, j);
                const double zc = sZ[a + 1][b + 1];
                const float axw = ax[k], axe = ax[k + G.sy], ays = ay[k], ayn = ay[k + 1];
                const double off = axw * static_cast<double>(sZ[a][b + 1]) + axe * static_cast<double>(sZ[a + 2][b + 1]) +
                                   ays * static_cast<double>(sZ[a + 1][b]) + ayn * static_cast<double>(sZ[a + 1][b + 2]);
                const float rf = sR[a][b];
                if (UPD) {
                    r_out[k] = rf;
                    x[k] = alpha * pu_dir[k] + x[k];
                    a_rr += static_cast<double>(rf) * rf;
                }
                rv = rf;
                w = MULT ? (1.0 - omega) * rv - omega * off : -off;
                a_rz += zc * rv;
#pragma unroll
                for (int jj = 0; jj < MW; jj++)
                    if (jj < nwin) {
                        qv[jj] = (UPD && jj == 0) ? sQ[UPD ? a : 0][b] : qj[jj][k];
                        a_zq[jj] += zc * qv[jj];
                    }
                if (MULT) {
                    const double as = rv + off;  // A D^-1 r (its diagonal term is r)
                    a_ss += as * zc;
                    if (mq) {
                        a_sq += as * static_cast<double>(sDQ[MULT ? a + 1 : 0][MULT ? b + 1 : 0]);
                        aqv = qv[0] + (axw * static_cast<double>(sDQ[MULT ? a : 0][MULT ? b + 1 : 0]) +
                                       axe * static_cast<double>(sDQ[MULT ? a + 2 : 0][MULT ? b + 1 : 0]) +
                                       ays * static_cast<double>(sDQ[MULT ? a + 1 : 0][MULT ? b : 0]) +
                                       ayn * static_cast<double>(sDQ[MULT ? a + 1 : 0][MULT ? b + 2 : 0]));
                    }
                }
            }
            cw += w;
            cr += rv;
#pragma unroll
            for (int jj = 0; jj < MW; jj++) cq[jj] += qv[jj];
            if (MULT) caq += aqv;
            if ((it_r + 1) % RPW != 0) continue;  // more rows of this block for this wave
            cw = seg_sum(cw, c);
            cr = seg_sum(cr, c);
#pragma unroll
            for (int jj = 0; jj < MW; jj++)
                if (jj < nwin) cq[jj] = seg_sum(cq[jj], c);
            if (mq) caq = seg_sum(caq, c);
            if ((lane & (c - 1)) == 0) {
                const int blk = (a / c) * nbj + b / c;
                if constexpr (NSLOT > 1) {
                    double (*sw)[nblk] = sB[threadIdx.x >> 6];  // this wave's slot
                    sw[0][blk] = cw;
                    sw[1][blk] = cr;
#pragma unroll
                    for (int jj = 0; jj < MW; jj++)
                        if (jj < nwin) sw[2 + jj][blk] = cq[jj];
                    if (mq) sw[NS - 1][blk] = caq;
                } else {
                    atomicAdd(&sB[0][0][blk], cw);
                    atomicAdd(&sB[0][1][blk], cr);
#pragma unroll
                    for (int jj = 0; jj < MW; jj++)
                        if (jj < nwin) atomicAdd(&sB[0][2 + jj][blk], cq[jj]);
                    if (mq) atomicAdd(&sB[0][NS - 1][blk], caq);
                }
            }
            cw = cr = caq = 0.0;
#pragma unroll
            for (int jj = 0; jj < MW; jj++) cq[jj] = 0.0;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nblk; e += KB) {
            const int bi = e / nbj, bj = e - bi * nbj;
            const int I = (i0 - ilo) / c + bi, J = (j0 - jlo) / c + bj;
            if (I < C.Hc && J < C.Wc) {
                const int64_t pc = static_cast<int64_t>(I) * C.Wc + J;
                auto slots = [&](int k) {
                    double t = sB[0][k][e];
                    for (int sl = 1; sl < NSLOT; sl++) t += sB[sl][k][e];
                    return t;
                };
                const double gv = slots(0);
                C.g[pc] = gv;
                C.rc[pc] = slots(1);
#pragma unroll
                for (int jj = 0; jj < MW; jj++)
                    if (jj < nwin) C.qc[jj][pc] = slots(2 + jj);
                if (mq) C.aq[pc] = slots(NS - 1);
                const double u = gv / C.dc[I * (C.Wc + 1) + J];
                a_uu += u * u;
            }
        }
        __syncthreads();
    }
    double v[2 + FCG_M] = {};  // the partial rows are laid out for FCG_M directions
    v[0] = a_rz;
#pragma unroll
    for (int jj = 0; jj < MW; jj++) v[1 + jj] = a_zq[jj];
    v[1 + FCG_M] = a_uu;
    put_partials(v, pb);
    if (MULT) {
        double m2[2] = {a_ss, a_sq};
        put_partials(m2, pm);
    }
    if (UPD) {
        double w2[2] = {a_rr, 0.0};
        put_partials(w2, pu);
        if (blockIdx.x == 0 && threadIdx.x == 0) st->iter = itn;
    }
    snapshot();
}

// linear-pass input channel 0 = g / dc / s, s = rms over the coarse cells of g / dc
__global__ __launch_bounds__(KB) void k_lp_pack(Coarse C, uint16_t* __restrict__ flin, const KrylovState* st,
                                                const double* pb, int nbb) {
    if (st->done != DONE_RUNNING) return;
    double v[1];
    sum_partials_all(pb + (1 + FCG_M) * nbb, nbb, v);  // row 1 + FCG_M: sum (g / dc)^2
    const int64_t Pc = static_cast<int64_t>(C.Hc) * C.Wc;
    double s = sqrt(v[0] / static_cast<double>(Pc));
    if (!(s > 0.0) || !isfinite(s)) s = 1.0;
    const float is = static_cast<float>(1.0 / s);
    FLAT_LOOP(p, Pc) {
        const int I = static_cast<int>(p / C.Wc), J = static_cast<int>(p - static_cast<int64_t>(I) * C.Wc);
        flin[p * 4] = f2bf(static_cast<float>(C.g[p] / C.dc[I * (C.Wc + 1) + J]) * is);
    }
}

// partials (y.g, y.A_c y, y.P^T r, y.P^T q_1 .. y.P^T q_FCG_M)
// (and, with pcm, y.P^T A D^-1 q_1 of the multiplicative fine level)
__global__ __launch_bounds__(KB) void k_lp_coarse_dots(Coarse C, const float* __restrict__ y, const KrylovState* st,
                                                       int nwin, double* pc, double* pcm = nullptr) {
    if (st->done != DONE_RUNNING) return;
    const int64_t Pc = static_cast<int64_t>(C.Hc) * C.Wc;
    double a[3 + FCG_M] = {}, am[1] = {};
    FLAT_LOOP(p, Pc) {
        const int I = static_cast<int>(p / C.Wc), J = static_cast<int>(p - static_cast<int64_t>(I) * C.Wc);
        const double yv = y[p];
        a[0] += yv * C.g[p];
        a[1] += yv * coarse_apply(C.acx, C.acy, C.dc, y, C.Hc, C.Wc, I, J);
        a[2] += yv * C.rc[p];
#pragma unroll
        for (int j = 0; j < FCG_M; j++)
            if (j < nwin) a[3 + j] += yv * C.qc[j][p];
        if (pcm && nwin >= 1) am[0] += yv * C.aq[p];
    }
    put_partials(a, pc);
    if (pcm) put_partials(am, pcm);
}

// launches k_lp_pass<UPD, c, MW, MULT> for the run-time coarsening c
template <bool UPD, int MW, bool MULT = false>
void launch_lp_pass(hipStream_t s, int nb, Geom G, Coarse C, const float* r_in, float* r_out, float* x, Dirs R,
                    const float* dinv, const float* ax, const float* ay, KrylovState* st, const double* ps,
                    int nbs, double* pu, int itn, int nwin, double* pb, KrylovState* snap, double omega = 0.0,
                    double* pm = nullptr) {
#define ASD2_LP_PASS(CV)                                                                                              \
    hipLaunchKernelGGL((k_lp_pass<UPD, CV, MW, MULT>), dim3(nb), dim3(KB), 0, s, G, C, r_in, r_out, x, R, dinv, ax,   \
                       ay, st, ps, nbs, pu, itn, nwin, pb, snap, omega, pm)
    switch (C.c) {
        case 1: ASD2_LP_PASS(1); break;
        case 2: ASD2_LP_PASS(2); break;
        case 4: ASD2_LP_PASS(4); break;
        case 8: ASD2_LP_PASS(8); break;
        case 16: ASD2_LP_PASS(16); break;
        default: ASD2_LP_PASS(32); break;
    }
#undef ASD2_LP_PASS
}

// The decision of iteration `it` (converged / pause for a reliable update / limit), then gamma,
// rho = z.r, the coefficients of the last nwin directions, p = z - sum_j beta_j p_j on a tile
// with halo (z = D^-1 r + gamma P y formed here), q = A p, partial p.q.  beta_j =
// (z.q_j) / (p_j.q_j): z made A-orthogonal to the kept directions (Notay's truncated flexible CG;
// with a preconditioner that is the same in every iteration and nwin >= 1 this is PCG), each
// dot split into its fine Jacobi part and its coarse part.  gamma: the energy-optimal step along
// P y, or gamma_fixed when positive (a y that already carries its scale).  rho <= 0 records
// BREAKDOWN; the host then falls back to Jacobi.  y == nullptr: Jacobi (gamma = 0).  The diagonal
// of A is taken as 1 / D^-1 from the D^-1 tile that z needs anyway (the FP32 operator of the inner
// iteration; the FP64 true residual uses the assembled diagonal).
template <int MW, int CC>
__global__ __launch_bounds__(KB) void k_lp_spmv(Geom G, Coarse C, const float* __restrict__ ax,
                                                const float* __restrict__ ay, const float* __restrict__ dres,
                                                int ring_only,
                                                const float* __restrict__ r, const float* __restrict__ dinv,
                                                const float* __restrict__ y, Dirs Dd, KrylovState* st,
                                                const double* pb, int nbb, const double* pc, int nbc, double* ps,
                                                int it, int nwin, double gamma_fixed, const double* pu, int nbu,
                                                const double* fine_totals) {
    if (st->done != DONE_RUNNING) return;
    constexpr int FI = 32;  // tile rows (FI x TJ tiles)
    // The tile operands (r, D^-1, y, the kept directions) of a block's first tile are loaded into
    // registers before the scalars are reduced, so that the reduction's latency overlaps them;
    // later tiles are loaded while the previous one computes.
    constexpr int NH = (FI + 2) * (TJ + 2), NE = (NH + KB - 1) / KB;
    const int ipad = G.nx + 2 * G.g, jpad = G.ny + 2 * G.g;
    const int ilo = G.g, ihi = G.g + G.nx, jlo = G.g, jhi = G.g + G.ny;
    const int tiles_i = ceil_div(G.nx, FI), ntiles = tiles_i * ceil_div(G.ny, TJ);
    int t0, t1;
    tile_range(ntiles, t0, t1);
    float ld_r[NE], ld_d[NE], ld_y[NE], ld_p[NE][MW];
    auto load = [&](int t) {
        const int bj = t / tiles_i, bi = t - bj * tiles_i;
        const int j0 = G.g + bj * TJ, i0 = G.g + bi * FI;
#pragma unroll
        for (int u = 0; u < NE; u++) {
            const int e = threadIdx.x + u * KB, rr_ = e / (TJ + 2), cc = e - rr_ * (TJ + 2);
            const int i = i0 - 1 + rr_, j = j0 - 1 + cc;
            ld_r[u] = ld_d[u] = ld_y[u] = 0.0f;
#pragma unroll
            for (int jj = 0; jj < MW; jj++) ld_p[u][jj] = 0.0f;
            if (e < NH && i < ipad && j < jpad && i >= ilo && i < ihi && j >= jlo && j < jhi) {
                const int64_t k = G.idx(i, j);
                ld_r[u] = r[k];
                ld_d[u] = dinv[k];
                if (y) ld_y[u] = y[static_cast<int64_t>((i - ilo) / CC) * C.Wc + (j - jlo) / CC];
#pragma unroll
                for (int jj = 0; jj < MW; jj++)
                    if (jj < nwin) ld_p[u][jj] = Dd.p[jj][k];
            }
        }
    };
    if (t0 < t1) load(t0);
    // every block re-sums the rows this iteration needs: r.r of the kernel that made r; r.D^-1 r
    // and the MW (D^-1 r).q_j of the fine pass; y.g, y.A_c y, y.P^T r and the MW y.P^T q_j.  The
    // fine rows may come summed already (fine_totals, by a kernel between the pass and this one,
    // in the same order), which leaves the coarse rows.
    double v[1 + (1 + MW) + (3 + MW)];
    if (fine_totals) {
        double w[3 + MW];
        sum_partials3<0, 0, 3 + MW>(nullptr, 0, nullptr, 0, y ? pc : nullptr, nbc, w);
#pragma unroll
        for (int k = 0; k < 2 + MW; k++) v[k] = fine_totals[k];
#pragma unroll
        for (int k = 0; k < 3 + MW; k++) v[2 + MW + k] = w[k];
    } else {
        sum_partials3<1, 1 + MW, 3 + MW>(pu, nbu, pb, nbb, y ? pc : nullptr, nbc, v);
    }
    {
        // the decision of iteration `it`
        const double rn = sqrt(v[0]), resid = rn / st->bnorm;
        int done = DONE_RUNNING;
        if (!isfinite(resid)) done = DONE_BREAKDOWN;
        else if (resid <= st->tol || rn < st->delta * st->rmax || it >= st->maxit) done = DONE_PAUSED;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->rr = v[0];
            st->iter = it;
            if (done != DONE_RUNNING) st->done = done;
        }
        if (done != DONE_RUNNING) return;
    }
    const double* vb = v + 1;
    const double* vc = v + 2 + MW;
    double gm = 0.0;
    if (y) {
        const double t = vc[0] / vc[1];
        if (gamma_fixed > 0.0) gm = gamma_fixed;
        else if (vc[1] > 0.0 && isfinite(t)) gm = t;
    }
    const double rho = vb[0] + gm * vc[2];
    float coef[MW];
    bool ok = isfinite(rho) && rho > 0.0;
#pragma unroll
    for (int j = 0; j < MW; j++) {
        double b = 0.0;
        if (j < nwin) {
            const double zq = vb[1 + j] + gm * vc[3 + j];
            b = -zq / st->pqv[(it - 1 - j) % FCG_SLOTS];
            ok = ok && isfinite(b);
        }
        coef[j] = static_cast<float>(b);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (ok) {
            st->rz[it & 1] = rho;
            st->beta = coef[0];
            st->omega = gm;
        } else {
            st->done = DONE_BREAKDOWN;
            st->iter = it;
        }
    }
    if (!ok) return;
    const float gmf = static_cast<float>(gm);
    const float* pj[MW];
#pragma unroll
    for (int j = 0; j < MW; j++) pj[j] = Dd.p[j];
    float* __restrict__ pnew = Dd.pnew;
    float* __restrict__ q = Dd.qnew;
    __shared__ float sP[FI + 2][TJ + 2];
    const int tj = threadIdx.x & (TJ - 1), ti = threadIdx.x / TJ;
    double acc = 0.0;
    for (int t = t0; t < t1; t++) {
        const int bj = t / tiles_i, bi = t - bj * tiles_i;
        const int j0 = G.g + bj * TJ, i0 = G.g + bi * FI;
#pragma unroll
        for (int u = 0; u < NE; u++) {
            const int e = threadIdx.x + u * KB, rr_ = e / (TJ + 2), cc = e - rr_ * (TJ + 2);
            if (e >= NH) continue;
            const int i = i0 - 1 + rr_, j = j0 - 1 + cc;
            float val = 0.0f;
            if (i < ipad && j < jpad && i >= ilo && i < ihi && j >= jlo && j < jhi) {
                val = ld_d[u] * ld_r[u];
                if (y) val += gmf * ld_y[u];
#pragma unroll
                for (int jj = 0; jj < MW; jj++)
                    if (jj < nwin) val += coef[jj] * ld_p[u][jj];
            }
            sP[rr_][cc] = val;
        }
        if (t + 1 < t1) load(t + 1);
        __syncthreads();
        for (int rr_ = ti; rr_ < FI; rr_ += KB / TJ) {
            const int i = i0 + rr_, j = j0 + tj;
            if (i >= ihi || j >= jhi) continue;
            const int64_t k = G.idx(i, j);
            const float pcv = sP[rr_ + 1][tj + 1];
            // flux form (solver.h); the remainder of an assembled system lives on the boundary ring
            const bool ring = i == ilo || i == ihi - 1 || j == jlo || j == jhi - 1;
            const float dv = (!ring_only || ring) ? dres[k] : 0.0f;
            const float qv = ax[k] * (sP[rr_][tj + 1] - pcv) + ax[k + G.sy] * (sP[rr_ + 2][tj + 1] - pcv) +
                             ay[k] * (sP[rr_ + 1][tj] - pcv) + ay[k + 1] * (sP[rr_ + 1][tj + 2] - pcv) + dv * pcv;
            pnew[k] = pcv;
            q[k] = qv;
            acc += static_cast<double>(pcv) * qv;
        }
        __syncthreads();
    }
    double pq[1] = {acc};
    put_partials(pq, ps);
}

// launches k_lp_spmv<MW, c> for the run-time coarsening c
template <int MW>
void launch_lp_spmv(hipStream_t s, int nb, Geom G, Coarse C, const float* ax, const float* ay, const float* dres,
                    int ring_only, const float* r, const float* dinv, const float* y, Dirs Dd, KrylovState* st,
                    const double* pb, int nbb, const double* pc, int nbc, double* ps, int it, int nwin,
                    double gamma_fixed, const double* pu, int nbu, const double* fine_totals) {
#define ASD2_LP_SPMV(CV)                                                                                              \
    hipLaunchKernelGGL((k_lp_spmv<MW, CV>), dim3(nb), dim3(KB), 0, s, G, C, ax, ay, dres, ring_only, r, dinv, y, Dd, \
                       st, pb, nbb, pc, nbc, ps, it, nwin, gamma_fixed, pu, nbu, fine_totals)
    switch (C.c) {
        case 1: ASD2_LP_SPMV(1); break;
        case 2: ASD2_LP_SPMV(2); break;
        case 4: ASD2_LP_SPMV(4); break;
        case 8: ASD2_LP_SPMV(8); break;
        case 16: ASD2_LP_SPMV(16); break;
        default: ASD2_LP_SPMV(32); break;
    }
#undef ASD2_LP_SPMV
}

// The multiplicative fine level's direction kernel (one kept direction, the multilevel solve).
// z is the fine level's damped Jacobi step on either side of the coarse correction:
//   z1 = omega D^-1 r + P y,   z = z1 + omega D^-1 (r - A z1),
// y the V-cycle of g = P^T (r - omega A D^-1 r) (k_lp_pass<.., MULT>).  z is formed on a tile with a
// halo of one cell (A z1 from a halo of two), then p = z + b p_1 and q = A p on the tile.  With
// s = D^-1 r, z = 2 omega s - omega^2 D^-1 A s + P y - omega D^-1 A P y, so its dots split into
//   z.r   = 2 omega s.r - omega^2 (A s).s + y.g       (y.P^T A s = y.(P^T r - g) / omega)
//   z.q_1 = 2 omega s.q_1 - omega^2 (A s).D^-1 q_1 + y.P^T q_1 - omega y.P^T A D^-1 q_1,
// fine rows from the fine pass (pb: s.r, s.q_1; pm: (A s).s, (A s).D^-1 q_1), coarse rows from the
// level-0 up kernel (pc: y.g, y.A_c y, y.P^T r, y.P^T q_1; pcm: y.P^T A D^-1 q_1), so b =
// -(z.q_1) / (p_1.q_1) is known before p is formed.  The decision and the recorded scalars are
// those of k_lp_spmv with gamma = 1.
template <int CC>
__global__ __launch_bounds__(KB) void k_lp_spmv_mult(Geom G, Coarse C, const float* __restrict__ ax,
                                                     const float* __restrict__ ay, const float* __restrict__ dres,
                                                     int ring_only, const float* __restrict__ r,
                                                     const float* __restrict__ dinv, const float* __restrict__ y,
                                                     Dirs Dd, KrylovState* st, const double* pb, const double* pm,
                                                     int nbb, const double* pc, const double* pcm, int nbc, double* ps,
                                                     int it, int nwin, double omega, const double* pu, int nbu,
                                                     const double* fine_totals) {
    if (st->done != DONE_RUNNING) return;
    // A tile of FI x FJ cells sits in an operand region of RH x RW = 32 x 64 cells (a halo of two):
    // lane = region column, wave w = region rows w, w + 4, ..., so every index is a lane or a row.
    // (40-row regions, two tiles per block at 2048^2, ran at 3 waves per SIMD and took longer.)
    constexpr int RH = 32, RW = 64, FI = RH - 4, FJ = RW - 4;
    constexpr int NWV = KB / 64, NR = RH / NWV;
    static_assert(RW == 64, "one region row per wave");
    const int ilo = G.g, ihi = G.g + G.nx, jlo = G.g, jhi = G.g + G.ny;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // Each block takes `per` consecutive tile slots of one column of tiles, the column padded to a
    // multiple of `per` (empty slots are skipped), so that the blocks of horizontally adjacent
    // columns (logical indices tiles_ip / per apart, on the same XCD) work on the same rows at the
    // same time and the cache lines their region edges share are fetched once.
    const int tiles_i = ceil_div(G.nx, FI), tiles_j = ceil_div(G.ny, FJ);
    int per = ceil_div(tiles_i * tiles_j, gridDim.x), tiles_ip = ceil_div(tiles_i, per) * per;
    while (ceil_div(tiles_ip * tiles_j, gridDim.x) > per) {
        per = ceil_div(tiles_ip * tiles_j, gridDim.x);
        tiles_ip = ceil_div(tiles_i, per) * per;
    }
    const int t0 = logical_block() * per;
    int t1 = min(tiles_ip * tiles_j, t0 + per);
    if (t0 < t1 && t0 % tiles_ip + (t1 - t0) > tiles_i) t1 = t0 + max(0, tiles_i - t0 % tiles_ip);  // empty slots
    const float* __restrict__ p1 = Dd.p[0];
    // A tile's operands in one batch of loads: r, D^-1, y, p_1 (0 outside the interior), the W-face
    // couplings (rows up to the ghost row below the interior) and the S-face couplings (columns up
    // to the ghost column), which hold the faces the tile and its ring use.
    // A block's consecutive tiles are vertically adjacent within a column of tiles: the next
    // region's rows 0 .. 3 are this region's rows FI .. FI + 3, which each wave holds as its last
    // row (FI is a multiple of the wave count), so they move from u = NR - 1 to u = 0 instead of
    // being loaded again.
    static_assert(FI % NWV == 0, "the carried rows stay with their waves");
    float ld_r[NR], ld_d[NR], ld_y[NR], ld_p[NR], ld_ax[NR], ld_ay[NR];
    auto load = [&](int t, bool carry) {
        const int bj = t / tiles_ip, bi = t - bj * tiles_ip;
        const int i0 = ilo + bi * FI - 2, j = jlo + bj * FJ - 2 + lane;  // the region's first row, this lane's column
        const bool jin = j >= jlo && j < jhi, jiny = j >= jlo && j <= jhi;
        const int jc = (j - jlo) / CC;
#pragma unroll
        for (int u = 0; u < NR; u++) {
            if (u == 0 && carry) {
                ld_r[0] = ld_r[NR - 1];
                ld_d[0] = ld_d[NR - 1];
                ld_y[0] = ld_y[NR - 1];
                ld_p[0] = ld_p[NR - 1];
                ld_ax[0] = ld_ax[NR - 1];
                ld_ay[0] = ld_ay[NR - 1];
                continue;
            }
            const int i = i0 + wv + NWV * u;
            const bool in = jin && i >= ilo && i < ihi;
            const bool inx = jin && i >= ilo && i <= ihi;
            const bool iny = jiny && i >= ilo && i < ihi;
            // 32-bit offsets (a padded array holds fewer than 2^31 cells)
            const unsigned k = static_cast<unsigned>(j + G.sy * i);
            const unsigned kc = in ? k : 0u;
            const float vr = r[kc], vd = dinv[kc];
            const float vy = y[in ? static_cast<unsigned>(((i - ilo) / CC) * C.Wc + jc) : 0u];
            const float vp = p1[(in && nwin >= 1) ? k : 0u];
            const float vx = ax[inx ? k : 0u], vya = ay[iny ? k : 0u];
            ld_r[u] = in ? vr : 0.0f;
            ld_d[u] = in ? vd : 0.0f;
            ld_y[u] = in ? vy : 0.0f;
            ld_p[u] = (in && nwin >= 1) ? vp : 0.0f;
            ld_ax[u] = inx ? vx : 0.0f;
            ld_ay[u] = iny ? vya : 0.0f;
        }
    };
    if (t0 < t1) load(t0, false);
    // f: r.r, s.r, s.q_1, (A s).s, (A s).D^-1 q_1; cv: y.g, y.A_c y, y.P^T r, y.P^T q_1, y.P^T A D^-1 q_1
    double f[5], cv[5];
    if (fine_totals) {
#pragma unroll
        for (int k = 0; k < 5; k++) f[k] = fine_totals[k];
    } else {
        sum_partials3<1, 2, 2>(pu, nbu, pb, nbb, pm, nbb, f);
    }
    sum_partials3<0, 4, 1>(nullptr, 0, pc, nbc, pcm, nbc, cv);
    {
        // the decision of iteration `it`
        const double rn = sqrt(f[0]), resid = rn / st->bnorm;
        int done = DONE_RUNNING;
        if (!isfinite(resid)) done = DONE_BREAKDOWN;
        else if (resid <= st->tol || rn < st->delta * st->rmax || it >= st->maxit) done = DONE_PAUSED;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->rr = f[0];
            st->iter = it;
            if (done != DONE_RUNNING) st->done = done;
        }
        if (done != DONE_RUNNING) return;
    }
    const double w2 = omega * omega;
    const double rho = 2.0 * omega * f[1] - w2 * f[3] + cv[0];
    double b = 0.0;
    bool ok = isfinite(rho) && rho > 0.0;
    if (nwin >= 1) {
        const double zq = 2.0 * omega * f[2] - w2 * f[4] + cv[3] - omega * cv[4];
        b = -zq / st->pqv[(it - 1) % FCG_SLOTS];
        ok = ok && isfinite(b);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (ok) {
            st->rz[it & 1] = rho;
            st->beta = b;
            st->omega = 1.0;
        } else {
            st->done = DONE_BREAKDOWN;
            st->iter = it;
        }
    }
    if (!ok) return;
    const float bf = static_cast<float>(b), omf = static_cast<float>(omega);
    float* __restrict__ pnew = Dd.pnew;
    float* __restrict__ q = Dd.qnew;
    __shared__ float sZ[RH][RW];   // z1 on the region
    __shared__ float sAX[RH][RW];  // W-face couplings
    __shared__ float sAY[RH][RW];  // S-face couplings
    __shared__ float sP[RH][RW];   // p on the tile and its ring
    double acc = 0.0;
    for (int t = t0; t < t1; t++) {
        const int bj = t / tiles_ip, bi = t - bj * tiles_ip;
        const int i0 = ilo + bi * FI - 2, j = jlo + bj * FJ - 2 + lane;
        const bool jin = j >= jlo && j < jhi;
#pragma unroll
        for (int u = 0; u < NR; u++) {
            const int rr_ = wv + NWV * u;
            sZ[rr_][lane] = omf * ld_d[u] * ld_r[u] + ld_y[u];  // 0 outside the interior (every operand 0)
            sAX[rr_][lane] = ld_ax[u];
            sAY[rr_][lane] = ld_ay[u];
        }
        __syncthreads();
        // z = z1 + omega D^-1 (r - A z1) and p = z + b p_1 on the tile and its ring (r, D^-1, p_1
        // from the registers)
#pragma unroll
        for (int u = 0; u < NR; u++) {
            const int rr_ = wv + NWV * u, i = i0 + rr_;
            if (rr_ < 1 || rr_ > RH - 2 || lane < 1 || lane > RW - 2) continue;
            float val = 0.0f;
            if (jin && i >= ilo && i < ihi) {
                const float zc = sZ[rr_][lane];
                const bool ring = i == ilo || i == ihi - 1 || j == jlo || j == jhi - 1;
                const float dv = (!ring_only || ring) ? dres[G.idx(i, j)] : 0.0f;
                const float az = sAX[rr_][lane] * (sZ[rr_ - 1][lane] - zc) + sAX[rr_ + 1][lane] * (sZ[rr_ + 1][lane] - zc) +
                                 sAY[rr_][lane] * (sZ[rr_][lane - 1] - zc) + sAY[rr_][lane + 1] * (sZ[rr_][lane + 1] - zc) +
                                 dv * zc;
                val = zc + omf * ld_d[u] * (ld_r[u] - az);
                if (nwin >= 1) val += bf * ld_p[u];
            }
            sP[rr_][lane] = val;
        }
        __syncthreads();
        if (t + 1 < t1) load(t + 1, true);  // a block's slots never cross a column of tiles
#pragma unroll
        for (int u = 0; u < NR; u++) {
            const int rr_ = wv + NWV * u, i = i0 + rr_;
            if (rr_ < 2 || rr_ > RH - 3 || lane < 2 || lane > RW - 3 || !jin || i >= ihi) continue;
            const int64_t k = G.idx(i, j);
            const float pcv = sP[rr_][lane];
            const bool ring = i == ilo || i == ihi - 1 || j == jlo || j == jhi - 1;
            const float dv = (!ring_only || ring) ? dres[k] : 0.0f;
            const float qv = sAX[rr_][lane] * (sP[rr_ - 1][lane] - pcv) + sAX[rr_ + 1][lane] * (sP[rr_ + 1][lane] - pcv) +
                             sAY[rr_][lane] * (sP[rr_][lane - 1] - pcv) + sAY[rr_][lane + 1] * (sP[rr_][lane + 1] - pcv) +
                             dv * pcv;
            pnew[k] = pcv;
            q[k] = qv;
            acc += static_cast<double>(pcv) * qv;
        }
        __syncthreads();
    }
    double pq[1] = {acc};
    put_partials(pq, ps);
}

// launches k_lp_spmv_mult<c> for the run-time coarsening c
void launch_lp_spmv_mult(hipStream_t s, int nb, Geom G, Coarse C, const float* ax, const float* ay, const float* dres,
                         int ring_only, const float* r, const float* dinv, const float* y, Dirs Dd, KrylovState* st,
                         const double* pb, const double* pm, int nbb, const double* pc, const double* pcm, int nbc,
                         double* ps, int it, int nwin, double omega, const double* pu, int nbu,
                         const double* fine_totals) {
#define ASD2_LP_SPMV_M(CV)                                                                                            \
    hipLaunchKernelGGL((k_lp_spmv_mult<CV>), dim3(nb), dim3(KB), 0, s, G, C, ax, ay, dres, ring_only, r, dinv, y, Dd, \
                       st, pb, pm, nbb, pc, pcm, nbc, ps, it, nwin, omega, pu, nbu, fine_totals)
    switch (C.c) {
        case 1: ASD2_LP_SPMV_M(1); break;
        case 2: ASD2_LP_SPMV_M(2); break;
        case 4: ASD2_LP_SPMV_M(4); break;
        case 8: ASD2_LP_SPMV_M(8); break;
        case 16: ASD2_LP_SPMV_M(16); break;
        default: ASD2_LP_SPMV_M(32); break;
    }
#undef ASD2_LP_SPMV_M
}

// r32 = r64 after a reliable update, partials (r.r, (D^-1 r)^2)
__global__ __launch_bounds__(KB) void k_fcg_resume(int64_t n, const double* __restrict__ r64, float* __restrict__ r32,
                                                   const float* __restrict__ dinv, const KrylovState* st,
                                                   double* partials) {
    if (st->done != DONE_PAUSED) return;
    double rr = 0.0, zz = 0.0;
    FLAT_LOOP(k, n) {
        const float rv = static_cast<float>(r64[k]);
        const float zv = dinv[k] * rv;
        r32[k] = rv;
        rr += static_cast<double>(rv) * rv;
        zz += static_cast<double>(zv) * zv;
    }
    double v[2] = {rr, zz};
    put_partials(v, partials);
}

// done: from -> to (single thread)
__global__ void k_set_done(KrylovState* st, int from, int to) {
    if (threadIdx.x == 0 && st->done == from) st->done = to;
}

// ------------------------------------------------------------------ multilevel coarse correction
// The same flexible PCG with y from a V-cycle instead of the network: the coarse operators of
// a hierarchy of block aggregations (level 0 = the c x c blocks of the fine grid, each next
// level factor 4 or 2 coarser, down to at most 16 x 16), Chebyshev smoothing before and after
// the coarse correction on every level, and the coarsest system solved exactly with its inverse,
// formed once per solve.  On the fine level a damped Jacobi step on either side of the coarse
// correction (k_lp_pass<.., MULT>, k_lp_spmv_mult) makes the preconditioner a symmetric
// multiplicative two-grid step; the step along P y is 1.  Levels of at least ML_WIDE cells run on
// the whole chip (tiles of T x T cells); the small tail of the hierarchy and the coarsest solve
// run in one workgroup.

constexpr int ML_MAX_LEVELS = 8;
constexpr int ML_MAX_DENSE = 256;   // unknowns of the coarsest level
constexpr int ML_T = 16;            // tile of the chip-wide level kernels on large levels
constexpr int ML_TS = 8;            // ... and on levels with fewer than 128 such tiles
constexpr int ML_SW = 4;            // smoothing steps before and after the coarse correction
// The smoother is the Chebyshev polynomial of degree ML_SW in D^-1 A on [LMAX / ALPHA, LMAX]
// (LMAX = 2 bounds the spectrum of D^-1 A for these diagonally dominant operators): step 0
// d = D^-1 r / theta, step k d = a_k d + b_k D^-1 r, each followed by y += d, r = g - A y.
constexpr double ML_LMAX = 2.0, ML_ALPHA = 32.0;
// damping of the fine level's Jacobi steps before and after the coarse correction
// (k_lp_spmv_mult): below 1 so that omega D^-1 A stays inside (0, 2) with a margin
constexpr double ML_OMEGA = 0.8;
struct Cheb {
    double c0;          // 1 / theta
    double a[ML_SW];    // a_k = rho_k rho_{k-1} (k >= 1)
    double b[ML_SW];    // b_k = 2 rho_k / delta (k >= 1)
};
Cheb ml_cheb() {
    Cheb c{};
    const double lo = ML_LMAX / ML_ALPHA, hi = ML_LMAX;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
    c.c0 = 1.0 / theta;
    for (int k = 1; k < ML_SW; k++) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        c.a[k] = rn * rho;
        c.b[k] = 2.0 * rn / delta;
        rho = rn;
    }
    return c;
}

struct Level {
    int H, W, c;              // c: factor to the next level (0 on the coarsest)
    double *acx, *acy, *dc;   // (H + 1) x (W + 1)
    double *g, *y1, *r1, *y;  // H x W
};

struct Hierarchy {
    int n;
    Level L[ML_MAX_LEVELS];
};

// operator of level `to` from level `from` (same aggregation rule as the first coarsening;
// the level arrays have the padded layout with no ghost layer and stride W + 1)
__global__ __launch_bounds__(KB) void k_ml_coarsen(Level from, Level to) {
    const int64_t cs = static_cast<int64_t>(to.H + 1) * (to.W + 1);
    FLAT_LOOP(e, cs) {
        const int I = static_cast<int>(e / (to.W + 1)), J = static_cast<int>(e - static_cast<int64_t>(I) * (to.W + 1));
        coarse_cell(from.acx, from.acy, from.dc, from.W + 1, 0, from.c, to.H, to.W, I, J, to.acx, to.acy, to.dc);
    }
}

// dense matrix of the coarsest level (row-major m x m, m = H W); one block
__global__ __launch_bounds__(KB) void k_ml_dense_build(Level L, double* __restrict__ M) {
    const int m = L.H * L.W;
    for (int64_t e = threadIdx.x; e < static_cast<int64_t>(m) * m; e += KB) M[e] = 0.0;
    __syncthreads();
    for (int p = threadIdx.x; p < m; p += KB) {
        const int I = p / L.W, J = p - I * L.W, k = I * (L.W + 1) + J;
        double* row = M + static_cast<int64_t>(p) * m;
        row[p] = L.dc[k];
        if (I > 0) row[p - L.W] = L.acx[k];
        if (I + 1 < L.H) row[p + L.W] = L.acx[k + L.W + 1];
        if (J > 0) row[p - 1] = L.acy[k];
        if (J + 1 < L.W) row[p + 1] = L.acy[k + 1];
    }
}

// ML_GJ_B consecutive Gauss-Jordan steps (pivots k0 .. k0 + ML_GJ_B - 1, no pivoting: the
// matrix is SPD) of the inverse of an m x m matrix, Mout from Min (two buffers, so no block
// reads a row another block rewrites); block = row i, thread = column j.  Every block replays
// the steps on the launch's pivot rows in registers, so after ceil(m / ML_GJ_B) launches the
// last output holds the inverse.  Per element each step computes what a one-step-per-launch
// sweep would: pivot row r_k / p, other rows r - f r_k / p, column k = 1 / p and -f / p.
constexpr int ML_GJ_B = 16;
__global__ __launch_bounds__(ML_MAX_DENSE) void k_ml_gj_block(const double* __restrict__ Min,
                                                              double* __restrict__ Mout, int m, int k0) {
    __shared__ double col[ML_GJ_B + 1];  // column k of the pivot rows and of the own row
    const int i = blockIdx.x, j = threadIdx.x;
    const bool act = j < m;
    const int nb = min(ML_GJ_B, m - k0);
    double P[ML_GJ_B];
#pragma unroll
    for (int t = 0; t < ML_GJ_B; t++)
        P[t] = (act && t < nb) ? Min[static_cast<int64_t>(k0 + t) * m + j] : 0.0;
    double x = act ? Min[static_cast<int64_t>(i) * m + j] : 0.0;
#pragma unroll
    for (int t = 0; t < ML_GJ_B; t++) {
        if (t >= nb) break;
        const int k = k0 + t;
        if (j == k) {
#pragma unroll
            for (int u = 0; u < ML_GJ_B; u++) col[u] = P[u];
            col[ML_GJ_B] = x;
        }
        __syncthreads();
        const double piv = 1.0 / col[t];
        const double rkj = P[t];
#pragma unroll
        for (int u = 0; u < ML_GJ_B; u++) {
            if (u == t) continue;
            const double f = col[u];
            P[u] = (j == k) ? -f * piv : P[u] - f * rkj * piv;
        }
        P[t] = (j == k) ? piv : rkj * piv;
        const double f = col[ML_GJ_B];
        x = (j == k) ? -f * piv : x - f * rkj * piv;
        __syncthreads();
    }
    const int own = i - k0;  // the own row is one of this launch's pivot rows
#pragma unroll
    for (int t = 0; t < ML_GJ_B; t++)
        if (t == own) x = P[t];
    if (act) Mout[static_cast<int64_t>(i) * m + j] = x;
}

// v of level cell (i, j), 0 outside the level grid
__device__ __forceinline__ double ml_at(const Level& L, const double* v, int i, int j) {
    return (i < 0 || i >= L.H || j < 0 || j >= L.W) ? 0.0 : v[static_cast<int64_t>(i) * L.W + j];
}
__device__ __forceinline__ double ml_apply(const Level& L, const double* v, int i, int j) {
    const int sc = L.W + 1, k = i * sc + j;
    return L.dc[k] * ml_at(L, v, i, j) + L.acx[k] * ml_at(L, v, i - 1, j) + L.acx[k + sc] * ml_at(L, v, i + 1, j) +
           L.acy[k] * ml_at(L, v, i, j - 1) + L.acy[k + 1] * ml_at(L, v, i, j + 1);
}

// A block's TH x TW window of a level array (row-major, leading dimension ld, nrows x ncols), top-left
// (ib, jb), 0 outside: every load issued before any value is used.
template <int TH, int TW, int NT>
__device__ __forceinline__ void lvl_fetch(const double* __restrict__ a, int ld, int nrows, int ncols,
                                          TileRegs<TH, TW, NT>& t, int ib, int jb) {
#pragma unroll
    for (int k = 0; k < TileRegs<TH, TW, NT>::NL; k++) {
        const int e = static_cast<int>(threadIdx.x) + k * NT;
        const int r = e / TW, c = e - r * TW;
        const int i = ib + r, j = jb + c;
        const bool in = e < TH * TW && i >= 0 && i < nrows && j >= 0 && j < ncols;
        const double x = a[in ? static_cast<int64_t>(i) * ld + j : 0];
        t.v[k] = in ? x : 0.0;
    }
}

// The operator and right-hand side of a level on a tile with the smoothing halo, staged in LDS:
// ax[a][b] = acx at region cell (a, b) (its W face; ax[a + 1][b] its E face), ay likewise along j,
// dc, g; the region's top-left is level cell (i0 - ML_SW, j0 - ML_SW).
template <int HW>
struct LevelTile {
    double ax[HW + 1][HW + 1], ay[HW + 1][HW + 1], dc[HW][HW], g[HW][HW];
};

template <int HW, int NT>
struct LevelTileRegs {
    TileRegs<HW + 1, HW + 1, NT> ax, ay;
    TileRegs<HW, HW, NT> dc, g;
    __device__ __forceinline__ void fetch(const Level& L, int i0, int j0) {
        const int sc = L.W + 1;
        lvl_fetch(L.acx, sc, L.H + 1, L.W + 1, ax, i0 - ML_SW, j0 - ML_SW);
        lvl_fetch(L.acy, sc, L.H + 1, L.W + 1, ay, i0 - ML_SW, j0 - ML_SW);
        lvl_fetch(L.dc, sc, L.H + 1, L.W + 1, dc, i0 - ML_SW, j0 - ML_SW);
        lvl_fetch(L.g, L.W, L.H, L.W, g, i0 - ML_SW, j0 - ML_SW);
    }
    __device__ __forceinline__ void store(LevelTile<HW>& s) const {
        tile_store(ax, s.ax);
        tile_store(ay, s.ay);
        tile_store(dc, s.dc);
        tile_store(g, s.g);
    }
};

// A y at region cell (a, b) from the staged operator (the expression of ml_apply / ml_step)
template <int HW>
__device__ __forceinline__ double ml_apply_s(const LevelTile<HW>& O, const double (*y)[HW], int a, int b) {
    return O.dc[a][b] * y[a][b] + O.ax[a][b] * y[a - 1][b] + O.ax[a + 1][b] * y[a + 1][b] + O.ay[a][b] * y[a][b - 1] +
           O.ay[a][b + 1] * y[a][b + 1];
}

// ml_step from the staged operator
template <int HW>
__device__ __forceinline__ double ml_step_s(const LevelTile<HW>& O, const Cheb& ch, int k, const double (*y)[HW],
                                            double& d, int a, int b) {
    const double ay = ml_apply_s<HW>(O, y, a, b);
    const double z = (O.g[a][b] - ay) / O.dc[a][b];
    d = (k == 0) ? ch.c0 * z : ch.a[k] * d + ch.b[k] * z;
    return y[a][b] + d;
}

// ml_steps from the staged operator
template <int T>
__device__ __forceinline__ void ml_steps_s(const Level& L, const LevelTile<T + 2 * ML_SW>& O, const Cheb& ch,
                                           double (*buf)[T + 2 * ML_SW][T + 2 * ML_SW],
                                           double (*dir)[T + 2 * ML_SW], int k0, int k1, int h0, int i0, int j0) {
    constexpr int HW = T + 2 * ML_SW;
    for (int k = k0; k < k1; k++) {
        const int s = k - k0;
        const double (*src)[HW] = buf[s & 1];
        double (*dst)[HW] = buf[(s + 1) & 1];
        const int lo = h0 + s + 1, n = HW - 2 * lo;  // region of halo ML_SW - lo
        for (int e = threadIdx.x; e < n * n; e += T * T) {
            const int a = lo + e / n, b = lo + e % n;
            const int i = i0 - ML_SW + a, j = j0 - ML_SW + b;
            dst[a][b] = (i >= 0 && i < L.H && j >= 0 && j < L.W) ? ml_step_s<HW>(O, ch, k, src, dir[a][b], a, b) : 0.0;
        }
        __syncthreads();
    }
}

// pre-smoothing and restriction on a level run by the whole chip: y1 = the ML_SW Chebyshev
// steps from 0, next.g = block sums of g - A y1.  A block of T x T threads owns a T x T tile
// (whole blocks of the next level: c divides T) and smooths it with a halo of ML_SW cells.  The
// tile's operator and g are staged in LDS in one batch of loads issued with the read of the solver
// state, so a launch waits for memory about once.
template <int T>
struct MlTileLds {
    static constexpr int HW = T + 2 * ML_SW;
    double y[2][HW][HW];
    double d[HW][HW];
    double r[T][T];
    LevelTile<HW> o;
};

// one tile of k_ml_down: false when the solve is not running (checked after the tile's loads are
// issued, so their latency overlaps the read of the state)
template <int T>
__device__ __forceinline__ bool ml_down_tile(const Level& L, const Level& next, const Cheb& ch, int done, int t,
                                             MlTileLds<T>& sm) {
    constexpr int HW = T + 2 * ML_SW, NT = T * T;
    const int tiles_j = ceil_div(L.W, T);
    const int c = L.c, nb = T / c;
    const int a = threadIdx.x / T, b = threadIdx.x - a * T;
    constexpr int F = (ML_SW - 1) & 1;
    const int ti = t / tiles_j, tj = t - ti * tiles_j;
    const int i0 = ti * T, j0 = tj * T;
    LevelTileRegs<HW, NT> tr;
    tr.fetch(L, i0, j0);
    if (done != DONE_RUNNING) return false;
    tr.store(sm.o);
    __syncthreads();
    for (int e = threadIdx.x; e < HW * HW; e += NT) {  // step 0 from y = 0: d = y = c0 g / dc
        const int aa = e / HW, bb = e - aa * HW;
        const int i = i0 - ML_SW + aa, j = j0 - ML_SW + bb;
        const double y0 = (i >= 0 && i < L.H && j >= 0 && j < L.W) ? ch.c0 * (sm.o.g[aa][bb] / sm.o.dc[aa][bb]) : 0.0;
        sm.y[0][aa][bb] = y0;
        sm.d[aa][bb] = y0;
    }
    __syncthreads();
    ml_steps_s<T>(L, sm.o, ch, sm.y, sm.d, 1, ML_SW, 0, i0, j0);
    const int i = i0 + a, j = j0 + b;
    double rv = 0.0;
    if (i < L.H && j < L.W) {
        const int A = a + ML_SW, B = b + ML_SW;
        const double (*y)[HW] = sm.y[F];
        const double ay = ml_apply_s<HW>(sm.o, y, A, B);
        L.y1[static_cast<int64_t>(i) * L.W + j] = y[A][B];
        rv = sm.o.g[A][B] - ay;
    }
    sm.r[a][b] = rv;
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < nb * nb) {
        const int bi = threadIdx.x / nb, bj = threadIdx.x - bi * nb;
        const int I = ti * nb + bi, J = tj * nb + bj;
        if (I < next.H && J < next.W) {
            double sum = 0.0;
            for (int u = 0; u < c; u++)
                for (int v = 0; v < c; v++) sum += sm.r[bi * c + u][bj * c + v];
            next.g[static_cast<int64_t>(I) * next.W + J] = sum;
        }
    }
    __syncthreads();
    return true;
}

// pre-smoothing and restriction on a level run by the whole chip: y1 = the ML_SW Chebyshev
// steps from 0, next.g = block sums of g - A y1.  A block of T x T threads owns a T x T tile
// (whole blocks of the next level: c divides T) and smooths it with a halo of ML_SW cells.  The
// tile's operator and g are staged in LDS in one batch of loads issued with the read of the solver
// state, so a launch waits for memory about once.
template <int T>
__global__ __launch_bounds__(T * T) void k_ml_down(Level L, Level next, const KrylovState* st, Cheb ch) {
    __shared__ MlTileLds<T> sm;
    const int done = st->done;
    const int ntiles = ceil_div(L.H, T) * ceil_div(L.W, T);
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x)
        if (!ml_down_tile<T>(L, next, ch, done, t, sm)) return;
}

// y = A^-1 g on the coarsest level with 4 threads per row (FP64), in a block of 1024
__device__ __forceinline__ void ml_dense(const Level& L, const double* __restrict__ Minv, double* __restrict__ sred) {
    const int m = L.H * L.W;
    const int row = threadIdx.x >> 2, part = threadIdx.x & 3;
    for (int p0 = 0; p0 < m; p0 += 256) {
        const int p = p0 + row;
        double s = 0.0;
        if (p < m) {
            const double* rw = Minv + static_cast<int64_t>(p) * m;
            for (int q = part; q < m; q += 4) s += rw[q] * L.g[q];
        }
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        if (p < m && part == 0) L.y[p] = s;
    }
    (void)sred;
}

// levels first .. n-1 in one workgroup (the small tail of the V-cycle; g of level `first`
// given), y of level `first`
__global__ __launch_bounds__(1024) void k_ml_lower(Hierarchy Hh, int first, const double* __restrict__ Minv,
                                                   const KrylovState* st, Cheb ch) {
    if (st->done != DONE_RUNNING) return;
    const int n = Hh.n;
    for (int l = first; l + 1 < n; l++) {
        const Level& L = Hh.L[l];
        const Level& N = Hh.L[l + 1];
        const int P = L.H * L.W;
        double* d = L.y;  // the Chebyshev directions (L.y is written only by the prolongation)
        for (int p = threadIdx.x; p < P; p += 1024) {  // step 0 from y = 0
            const int i = p / L.W, j = p - i * L.W;
            const double y0 = ch.c0 * (L.g[p] / L.dc[i * (L.W + 1) + j]);
            L.y1[p] = y0;
            d[p] = y0;
        }
        __syncthreads();
        for (int k = 1; k < ML_SW; k++) {  // further steps through r1
            for (int p = threadIdx.x; p < P; p += 1024) {
                const int i = p / L.W, j = p - i * L.W;
                const double z = (L.g[p] - ml_apply(L, L.y1, i, j)) / L.dc[i * (L.W + 1) + j];
                d[p] = ch.a[k] * d[p] + ch.b[k] * z;
                L.r1[p] = L.y1[p] + d[p];
            }
            __syncthreads();
            for (int p = threadIdx.x; p < P; p += 1024) L.y1[p] = L.r1[p];
            __syncthreads();
        }
        for (int p = threadIdx.x; p < P; p += 1024) {
            const int i = p / L.W, j = p - i * L.W;
            L.r1[p] = L.g[p] - ml_apply(L, L.y1, i, j);
        }
        __syncthreads();
        for (int p = threadIdx.x; p < N.H * N.W; p += 1024) {
            const int I = p / N.W, J = p - I * N.W;
            double s = 0.0;
            for (int u = 0; u < L.c; u++)
                for (int v = 0; v < L.c; v++) s += L.r1[static_cast<int64_t>(I * L.c + u) * L.W + J * L.c + v];
            N.g[p] = s;
        }
        __syncthreads();
    }
    ml_dense(Hh.L[n - 1], Minv, nullptr);
    __syncthreads();
    for (int l = n - 2; l >= first; l--) {
        const Level& L = Hh.L[l];
        const Level& N = Hh.L[l + 1];
        const int P = L.H * L.W;
        for (int p = threadIdx.x; p < P; p += 1024) {  // y2 into r1
            const int i = p / L.W, j = p - i * L.W;
            L.r1[p] = L.y1[p] + N.y[static_cast<int64_t>(i / L.c) * N.W + j / L.c];
        }
        __syncthreads();
        double* d = L.y1;  // the Chebyshev directions (y1 is folded into r1)
        for (int s = 0; s < ML_SW; s++) {
            for (int p = threadIdx.x; p < P; p += 1024) {
                const int i = p / L.W, j = p - i * L.W;
                const double z = (L.g[p] - ml_apply(L, L.r1, i, j)) / L.dc[i * (L.W + 1) + j];
                d[p] = (s == 0) ? ch.c0 * z : ch.a[s] * d[p] + ch.b[s] * z;
                L.y[p] = L.r1[p] + d[p];
            }
            __syncthreads();
            if (s + 1 < ML_SW) {
                for (int p = threadIdx.x; p < P; p += 1024) L.r1[p] = L.y[p];
                __syncthreads();
            }
        }
    }
}

// prolongation and post-smoothing on a level run by the whole chip: y = the ML_SW Chebyshev
// steps from y2 = y1 + P y_next (on T x T tiles with the sweeps' halo), into L.y, or into yout
// (FP32) on level 0, where the block partials (0, 0, y.P^T r, y.P^T q_1) of the fine
// iteration's scalars go to pc (the step along P y is fixed at 1, so y.g and y.A_c y are not
// needed).  The tile's operator, y1, the coarse values under it and (level 0) the coarse dots'
// operands are loaded in one batch with the read of the solver state.
// one tile of k_ml_up (yout, rc, qc, aqc: level 0 only; a_r, a_q, a_g, a_aq accumulate the partials'
// operands y.P^T r, y.P^T q_1, and with aqc the multiplicative level's y.g and y.P^T A D^-1 q_1)
template <int T>
__device__ __forceinline__ bool ml_up_tile(const Level& L, const Level& next, float* __restrict__ yout,
                                           const double* rc, const double* qc, const double* aqc, const Cheb& ch,
                                           int done, int t, MlTileLds<T>& sm, double& a_r, double& a_q, double& a_g,
                                           double& a_aq) {
    constexpr int HW = T + 2 * ML_SW, NT = T * T;
    const int tiles_j = ceil_div(L.W, T);
    const int a = threadIdx.x / T, b = threadIdx.x - a * T;
    constexpr int F = (ML_SW - 1) & 1;
    const int ti = t / tiles_j, tj = t - ti * tiles_j;
    const int i0 = ti * T, j0 = tj * T;
    LevelTileRegs<HW, NT> tr;
    tr.fetch(L, i0, j0);
    TileRegs<HW, HW, NT> ty1, tyn;
    lvl_fetch(L.y1, L.W, L.H, L.W, ty1, i0 - ML_SW, j0 - ML_SW);
#pragma unroll
    for (int k = 0; k < TileRegs<HW, HW, NT>::NL; k++) {  // P y_next at the region's cells
        const int e = static_cast<int>(threadIdx.x) + k * NT;
        const int aa = e / HW, bb = e - aa * HW;
        const int i = i0 - ML_SW + aa, j = j0 - ML_SW + bb;
        const bool in = e < HW * HW && i >= 0 && i < L.H && j >= 0 && j < L.W;
        const double x = next.y[in ? static_cast<int64_t>(i / L.c) * next.W + j / L.c : 0];
        tyn.v[k] = in ? x : 0.0;
    }
    const int i = i0 + a, j = j0 + b;
    const bool mine = i < L.H && j < L.W;
    const int64_t e0 = mine ? static_cast<int64_t>(i) * L.W + j : 0;
    const double rcv = yout ? rc[e0] : 0.0, qcv = yout ? qc[e0] : 0.0;
    const double gcv = (yout && aqc) ? L.g[e0] : 0.0, aqv = (yout && aqc) ? aqc[e0] : 0.0;
    if (done != DONE_RUNNING) return false;
    tr.store(sm.o);
#pragma unroll
    for (int k = 0; k < TileRegs<HW, HW, NT>::NL; k++) {  // y2 = y1 + P y_next (0 outside the level)
        const int e = static_cast<int>(threadIdx.x) + k * NT;
        if ((HW * HW) % NT == 0 || e < HW * HW) {
            const int aa = e / HW, bb = e - aa * HW;
            sm.y[0][aa][bb] = ty1.v[k] + tyn.v[k];
        }
    }
    __syncthreads();
    ml_steps_s<T>(L, sm.o, ch, sm.y, sm.d, 0, ML_SW - 1, 0, i0, j0);
    if (mine) {
        const double y = ml_step_s<HW>(sm.o, ch, ML_SW - 1, sm.y[F], sm.d[a + ML_SW][b + ML_SW], a + ML_SW, b + ML_SW);
        if (yout) {
            const float yf = static_cast<float>(y);
            yout[e0] = yf;
            a_r += static_cast<double>(yf) * rcv;
            a_q += static_cast<double>(yf) * qcv;
            a_g += static_cast<double>(yf) * gcv;
            a_aq += static_cast<double>(yf) * aqv;
        } else {
            L.y[e0] = y;
        }
    }
    __syncthreads();
    return true;
}

// prolongation and post-smoothing on a level run by the whole chip: y = the ML_SW Chebyshev
// steps from y2 = y1 + P y_next (on T x T tiles with the sweeps' halo), into L.y, or into yout
// (FP32) on level 0, where the block partials (0, 0, y.P^T r, y.P^T q_1) of the fine
// iteration's scalars go to pc (the step along P y is fixed at 1, so y.g and y.A_c y are not
// needed).  The tile's operator, y1, the coarse values under it and (level 0) the coarse dots'
// operands are loaded in one batch with the read of the solver state.
// With aqc (the multiplicative fine level) the level-0 launch also writes y.g into row 0 of pc and
// y.P^T A D^-1 q_1 into pcm.
template <int T>
__global__ __launch_bounds__(T * T) void k_ml_up(Level L, Level next, float* __restrict__ yout, const double* rc,
                                                 const double* qc, const KrylovState* st, double* pc, Cheb ch,
                                                 const double* aqc = nullptr, double* pcm = nullptr) {
    __shared__ MlTileLds<T> sm;
    const int done = st->done;
    const int ntiles = ceil_div(L.H, T) * ceil_div(L.W, T);
    double a_r = 0.0, a_q = 0.0, a_g = 0.0, a_aq = 0.0;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x)
        if (!ml_up_tile<T>(L, next, yout, rc, qc, aqc, ch, done, t, sm, a_r, a_q, a_g, a_aq)) return;
    if constexpr (T * T == KB) {  // the partials' block reduction assumes KB threads (level 0 tiles)
        if (pc) {
            double v[4] = {a_g, 0.0, a_r, a_q};
            put_partials(v, pc);
        }
        if (pcm) {
            double v[1] = {a_aq};
            put_partials(v, pcm);
        }
    }
}

// the fine pass's scalar rows for the SpMV (r.r, r.D^-1 r, (D^-1 r).q_1), summed once in the order
// every SpMV block would use
struct FineRows {
    const double* pu;
    int nbu;
    const double* pb;
    int nbb;
    double* out;  // nullptr: not summed here
    const double* pm = nullptr;  // the multiplicative fine level's two rows, summed after the three
};

// y = A^-1 g on the coarsest level (FP64 inverse): one row per wave, the lanes splitting it; with
// F.out set, one more block sums the fine rows (the pass that wrote them has finished)
__global__ __launch_bounds__(KB) void k_ml_dense_apply(Level L, const double* __restrict__ Minv,
                                                       const KrylovState* st, FineRows F) {
    if (st->done != DONE_RUNNING) return;
    if (F.out && blockIdx.x == gridDim.x - 1) {
        if (F.pm) {
            double v[5];
            sum_partials3<1, 2, 2>(F.pu, F.nbu, F.pb, F.nbb, F.pm, F.nbb, v);
            if (threadIdx.x == 0)
                for (int k = 0; k < 5; k++) F.out[k] = v[k];
        } else {
            double v[3];
            sum_partials3<1, 2, 0>(F.pu, F.nbu, F.pb, F.nbb, nullptr, 0, v);
            if (threadIdx.x == 0)
                for (int k = 0; k < 3; k++) F.out[k] = v[k];
        }
        return;
    }
    const int m = L.H * L.W, lane = threadIdx.x & 63;
    const int p = blockIdx.x * (KB / 64) + (threadIdx.x >> 6);  // one row per wave
    if (p >= m) return;
    const double* row = Minv + static_cast<int64_t>(p) * m;
    double s = 0.0;
    for (int q = lane; q < m; q += 64) s += row[q] * L.g[q];
    s = wave_sum(s);
    if (lane == 0) L.y[p] = s;
}

// a single level (the first coarsening already small): y = A^-1 g, FP32 output; one block
__global__ __launch_bounds__(KB) void k_ml_dense_only(Level L, const double* __restrict__ Minv, float* __restrict__ yout,
                                                      const KrylovState* st) {
    if (st->done != DONE_RUNNING) return;
    const int m = L.H * L.W;
    for (int p = threadIdx.x; p < m; p += KB) {
        const double* row = Minv + static_cast<int64_t>(p) * m;
        double s = 0.0;
        for (int q = 0; q < m; q++) s += row[q] * L.g[q];
        yout[p] = static_cast<float>(s);
    }
}

// The learned coarse solve on level L of the hierarchy.  Network input channel 0 = g / dc: the
// linear pass has no biases, so its output scales with its input, and the step gamma below
// removes any scale; the other channels stay zero.
__global__ __launch_bounds__(KB) void k_ml_net_pack(Level L, uint16_t* __restrict__ flin, const KrylovState* st) {
    if (st->done != DONE_RUNNING) return;
    const int P = L.H * L.W, sc = L.W + 1;
    for (int p = blockIdx.x * KB + threadIdx.x; p < P; p += gridDim.x * KB) {
        const int i = p / L.W, j = p - i * L.W;
        flin[static_cast<int64_t>(p) * 4] = f2bf(static_cast<float>(L.g[p] / L.dc[i * sc + j]));
    }
}

// block partials of y.g and y.A_L y of the network's output y on level L
__global__ __launch_bounds__(KB) void k_ml_net_dots(Level L, const float* __restrict__ y, const KrylovState* st,
                                                    double* part) {
    if (st->done != DONE_RUNNING) return;
    const int P = L.H * L.W;
    double v[2] = {0.0, 0.0};
    for (int p = blockIdx.x * KB + threadIdx.x; p < P; p += gridDim.x * KB) {
        const int i = p / L.W, j = p - i * L.W;
        const double yv = y[p];
        v[0] += yv * L.g[p];
        v[1] += yv * coarse_apply<double, float>(L.acx, L.acy, L.dc, y, L.H, L.W, i, j);
    }
    put_partials(v, part);
}

// L.y = gamma y, gamma = (y.g) / (y.A_L y): the energy-optimal step along the network's coarse
// correction on level L (of either sign; 0 when y.A_L y is not positive); every block re-sums the
// partials
__global__ __launch_bounds__(KB) void k_ml_net_apply(Level L, const float* __restrict__ y, const KrylovState* st,
                                                     const double* part, int nbp) {
    if (st->done != DONE_RUNNING) return;
    double v[2];
    sum_partials_all(part, nbp, v);
    const double t = v[0] / v[1];
    const double gm = (v[1] > 0.0 && isfinite(t)) ? t : 0.0;  // either sign: the loss does not fix y's
    const int P = L.H * L.W;
    for (int p = blockIdx.x * KB + threadIdx.x; p < P; p += gridDim.x * KB) L.y[p] = gm * static_cast<double>(y[p]);
}

// the hierarchy below a first coarsening of H0 x W0 cells
int ml_plan(int H0, int W0, int* H, int* W, int* c, int* nlev) {
    int n = 0;
    H[0] = H0;
    W[0] = W0;
    while ((H[n] > 16 || W[n] > 16) && H[n] % 2 == 0 && W[n] % 2 == 0 && n + 1 < ML_MAX_LEVELS) {
        const int f = (H[n] % 4 == 0 && W[n] % 4 == 0 && H[n] >= 8 && W[n] >= 8) ? 4 : 2;
        c[n] = f;
        H[n + 1] = H[n] / f;
        W[n + 1] = W[n] / f;
        n++;
    }
    c[n] = 0;
    *nlev = n + 1;
    if (H[n] * W[n] > ML_MAX_DENSE) {
        set_error("multilevel: the coarsest level (%d x %d) exceeds %d unknowns", H[n], W[n], ML_MAX_DENSE);
        return ASD2_ERR_UNSUPPORTED;
    }
    return ASD2_OK;
}

// 8 x 8 tiles on a level below the first where 16 x 16 tiles would give fewer than 128 blocks
// (level 0 keeps 16 x 16: its up kernel writes the fine iteration's partials, nbc of them)
bool ml_small_tiles(const Level& L, int l) {
    return l > 0 && ceil_div(L.H, ML_T) * ceil_div(L.W, ML_T) < 128 && L.c > 0 && ML_TS % L.c == 0;
}

// levels run by the whole chip (down/up kernels): the first ones down to fewer than
// ML_WIDE cells; the one-workgroup kernel takes the rest.  Level 0 always runs wide when there
// is a level below it: its output is the FP32 y of the fine iteration.
constexpr int ML_WIDE = 2048;
int ml_first_lower(const Hierarchy& Hh) {
    int f = Hh.n > 1 ? 1 : 0;
    while (f + 1 < Hh.n && Hh.L[f].H * Hh.L[f].W >= ML_WIDE) f++;
    return f;
}

// Flexible PCG with a coarse correction: learned (the network's linear form) or multilevel
// (a V-cycle on Galerkin aggregations).  Shared: the fine begin pass, the coarse dots, the
// direction/SpMV and update kernels, FP64 reliable updates and the Jacobi fallback.
int coarse_fcg_solve(asd2_ctx* ctx, const asd2_solve_opts* o, double* d_ps, asd2_solve_stats* stats, bool learned) {
    asd2_unet* N = ctx->net;
    const Geom& G = ctx->geo;
    int c0 = 1;
    if (learned) {
        if (!N) {
            set_error("no network attached to the context (asd2_set_preconditioner)");
            return ASD2_ERR_UNSUPPORTED;
        }
        if (N->H * N->coarsen != G.nx || N->W * N->coarsen != G.ny) {
            set_error("the network was built for %d x %d, the grid is %d x %d", N->H * N->coarsen, N->W * N->coarsen,
                      G.nx, G.ny);
            return ASD2_ERR_UNSUPPORTED;
        }
    }
    while (c0 < 8 && G.nx % (2 * c0) == 0 && G.ny % (2 * c0) == 0) c0 *= 2;
    // The network replaces the bottom of the V-cycle: it solves on the hierarchy's level NL whose
    // coarsening equals the network's, the levels above it keep their Chebyshev smoothing.  A
    // network coarsening the hierarchy does not reach (or c0 itself) makes its own first level.
    int NL = 0;
    if (learned && N->coarsen != c0) {
        int Hl[ML_MAX_LEVELS], Wl[ML_MAX_LEVELS], cl[ML_MAX_LEVELS], nl = 0;
        bool found = false;
        if (N->coarsen > c0 && ml_plan(G.nx / c0, G.ny / c0, Hl, Wl, cl, &nl) == ASD2_OK) {
            int cum = c0;
            for (int l = 1; l < nl && !found; l++) {
                cum *= cl[l - 1];
                if (cum == N->coarsen) {
                    NL = l;
                    found = true;
                }
            }
        }
        if (!found) c0 = N->coarsen;
    }
    const bool hybrid = learned && NL > 0;
    double* net_part = nullptr;  // block partials of the network level's gamma
    int nbn = 1;
    const int64_t n = G.size();
    System S;
    ASD2_TRY(system_arrays(ctx, &S, false));
    ASD2_TRY(ensure_system32(ctx, &S));
    double *x64, *r64, *lp_part;
    float *x, *rb[2];  // r alternates: iteration it reads rb[it & 1]
    Ring R{};
    ASD2_TRY(ws(ctx, "cg_x", n, &x64));
    ASD2_TRY(ws(ctx, "cg_r", n, &r64));
    ASD2_TRY(ws(ctx, "mx_x", n, &x));
    ASD2_TRY(ws(ctx, "mx_r", n, &rb[0]));
    ASD2_TRY(ws(ctx, "mx_r2", n, &rb[1]));
    static const char* ring_names[2][FCG_SLOTS] = {{"fcg_p0", "fcg_p1", "fcg_p2", "fcg_p3", "fcg_p4"},
                                                   {"fcg_q0", "fcg_q1", "fcg_q2", "fcg_q3", "fcg_q4"}};
    for (int k = 0; k < FCG_SLOTS; k++) {
        ASD2_TRY(ws(ctx, ring_names[0][k], n, &R.p[k]));
        ASD2_TRY(ws(ctx, ring_names[1][k], n, &R.q[k]));
    }
    ASD2_TRY(ws(ctx, "lp_partials", (8 + 2 * FCG_M) * MAX_BLOCKS, &lp_part));
    Coarse C{};
    C.c = c0;
    C.Hc = G.nx / c0;
    C.Wc = G.ny / c0;
    const int64_t cs = static_cast<int64_t>(C.Hc + 1) * (C.Wc + 1), Pc = static_cast<int64_t>(C.Hc) * C.Wc;
    ASD2_TRY(ws(ctx, "lp_acx", cs, &C.acx));
    ASD2_TRY(ws(ctx, "lp_acy", cs, &C.acy));
    ASD2_TRY(ws(ctx, "lp_dc", cs, &C.dc));
    ASD2_TRY(ws(ctx, "lp_g", Pc, &C.g));
    ASD2_TRY(ws(ctx, "lp_rc", Pc, &C.rc));
    static const char* qc_names[FCG_M] = {"lp_qc0", "lp_qc1", "lp_qc2", "lp_qc3"};
    for (int j = 0; j < FCG_M; j++) ASD2_TRY(ws(ctx, qc_names[j], Pc, &C.qc[j]));
    ASD2_TRY(ws(ctx, "lp_aq", Pc, &C.aq));
    // multilevel hierarchy: level 0 shares the first coarsening's arrays
    Hierarchy Hh{};
    double *Minv = nullptr, *Mwork = nullptr;
    float* yml = nullptr;  // the V-cycle's output on level 0 (FP32, read by the SpMV)
    if (!learned || hybrid) {
        int Hl[ML_MAX_LEVELS], Wl[ML_MAX_LEVELS], cl[ML_MAX_LEVELS];
        ASD2_TRY(ml_plan(C.Hc, C.Wc, Hl, Wl, cl, &Hh.n));
        if (hybrid) {
            Hh.n = NL + 1;  // the network's level is the last one
            nbn = std::max(1, std::min(64, ceil_div(static_cast<int64_t>(Hl[NL]) * Wl[NL], KB)));
            ASD2_TRY(ws(ctx, "ml_net_part", 2 * nbn, &net_part));
        }
        static const char* names[ML_MAX_LEVELS][7] = {
            {"ml0_acx", "ml0_acy", "ml0_dc", "ml0_g", "ml0_y1", "ml0_r1", "ml0_y"},
            {"ml1_acx", "ml1_acy", "ml1_dc", "ml1_g", "ml1_y1", "ml1_r1", "ml1_y"},
            {"ml2_acx", "ml2_acy", "ml2_dc", "ml2_g", "ml2_y1", "ml2_r1", "ml2_y"},
            {"ml3_acx", "ml3_acy", "ml3_dc", "ml3_g", "ml3_y1", "ml3_r1", "ml3_y"},
            {"ml4_acx", "ml4_acy", "ml4_dc", "ml4_g", "ml4_y1", "ml4_r1", "ml4_y"},
            {"ml5_acx", "ml5_acy", "ml5_dc", "ml5_g", "ml5_y1", "ml5_r1", "ml5_y"},
            {"ml6_acx", "ml6_acy", "ml6_dc", "ml6_g", "ml6_y1", "ml6_r1", "ml6_y"},
            {"ml7_acx", "ml7_acy", "ml7_dc", "ml7_g", "ml7_y1", "ml7_r1", "ml7_y"}};
        for (int l = 0; l < Hh.n; l++) {
            Level& L = Hh.L[l];
            L.H = Hl[l];
            L.W = Wl[l];
            L.c = cl[l];
            const int64_t lc = static_cast<int64_t>(L.H + 1) * (L.W + 1), lp = static_cast<int64_t>(L.H) * L.W;
            if (l == 0) {
                L.acx = C.acx;
                L.acy = C.acy;
                L.dc = C.dc;
                L.g = C.g;
            } else {
                ASD2_TRY(ws(ctx, names[l][0], lc, &L.acx));
                ASD2_TRY(ws(ctx, names[l][1], lc, &L.acy));
                ASD2_TRY(ws(ctx, names[l][2], lc, &L.dc));
                ASD2_TRY(ws(ctx, names[l][3], lp, &L.g));
            }
            ASD2_TRY(ws(ctx, names[l][5], lp, &L.r1));
            ASD2_TRY(ws(ctx, names[l][4], lp, &L.y1));
            ASD2_TRY(ws(ctx, names[l][6], lp, &L.y));
        }
        if (!learned) {
            const int m = Hh.L[Hh.n - 1].H * Hh.L[Hh.n - 1].W;
            ASD2_TRY(ws(ctx, "ml_dense", static_cast<int64_t>(m) * m, &Minv));
            ASD2_TRY(ws(ctx, "ml_dense_work", static_cast<int64_t>(m) * m, &Mwork));
        }
        ASD2_TRY(ws(ctx, "ml_yout", Pc, &yml));
    }
    double* fine_totals = nullptr;  // the fine rows of the SpMV, summed by the dense-apply launch
    ASD2_TRY(ws(ctx, "lp_fine_totals", 6, &fine_totals));
    double* pb = lp_part;  // fine pass: r.D^-1 r, (D^-1 r).q_j, sum (g / dc)^2
    double* pc = lp_part + (2 + FCG_M) * MAX_BLOCKS;  // coarse dots: y.g, y.A_c y, y.P^T r, y.P^T q_j
    double* pm = lp_part + (5 + 2 * FCG_M) * MAX_BLOCKS;  // multiplicative level: (A s).s, (A s).D^-1 q_1
    double* pcm = pm + 2 * MAX_BLOCKS;                    // ... and y.P^T A D^-1 q_1
    const int nbs = spmv_blocks(G), nbf = flat_blocks(n), nbi = interior_blocks(G);
    const int RI = lp_ri(C.c);
    const int nbb = std::max(1, std::min<int>(MAX_BLOCKS, ceil_div(G.nx, RI) * ceil_div(G.ny, LP_RJ)));
    const int nbc = flat_blocks(Pc);
    asd2_solve_opts oo = *o;
    oo.mode = ASD2_CG_STANDARD;
    Driver D;
    ASD2_TRY(driver_init(ctx, &D));
    // the multilevel solve takes under a hundred iterations of one fixed preconditioner: its FP32
    // recurrence stays close to the true residual over four decades (1e-4: the same iterations
    // as 1e-2 at 2048^2 with half the updates; 1e-6 drifts and takes more), and every update
    // costs a host round trip, so its default factor is 1e-4; the learned solve keeps 0.1
    ASD2_TRY(set_state(&D, &oo, learned ? 0.1 : 1e-4));
    hipStream_t saved = learned ? N->stream : nullptr;
    if (learned) N->stream = ctx->stream;
    struct Restore {
        asd2_unet* N;
        hipStream_t s;
        ~Restore() {
            if (N) N->stream = s;
        }
    } restore{learned ? N : nullptr, saved};
    hipLaunchKernelGGL(k_lp_coarse_op, dim3(flat_blocks(cs)), dim3(KB), 0, ctx->stream, G, S.ax, S.ay, S.diag, C);
    for (int l = 1; l < Hh.n; l++)
        hipLaunchKernelGGL(k_ml_coarsen, dim3(flat_blocks(static_cast<int64_t>(Hh.L[l].H + 1) * (Hh.L[l].W + 1))),
                           dim3(KB), 0, ctx->stream, Hh.L[l - 1], Hh.L[l]);
    if (learned) {
        // the network's level as a coarse-grid view for the static features
        Coarse CN = C;
        if (hybrid) {
            CN.Hc = Hh.L[NL].H;
            CN.Wc = Hh.L[NL].W;
            CN.acx = Hh.L[NL].acx;
            CN.acy = Hh.L[NL].acy;
            CN.dc = Hh.L[NL].dc;
        }
        hipLaunchKernelGGL(k_lp_static_features, dim3(flat_blocks(static_cast<int64_t>(CN.Hc) * CN.Wc)), dim3(KB), 0,
                           ctx->stream, CN, N->feat, N->flin, static_cast<float>(G.dx * G.dx / N->coarsen),
                           static_cast<float>(G.dy * G.dy / N->coarsen));
        ASD2_TRY(unet_forward(N, 1));  // the ReLU pattern of this system, fixed for the solve
    } else {
        const Level& Lc = Hh.L[Hh.n - 1];
        const int m = Lc.H * Lc.W;
        // nl launches alternate A -> B -> A ...; the matrix is built where the last one ends in Minv
        const int nl = ceil_div(m, ML_GJ_B);
        double* A0 = (nl & 1) ? Mwork : Minv;
        double* B0 = (nl & 1) ? Minv : Mwork;
        hipLaunchKernelGGL(k_ml_dense_build, dim3(1), dim3(KB), 0, ctx->stream, Lc, A0);
        for (int q = 0; q < nl; q++)
            hipLaunchKernelGGL(k_ml_gj_block, dim3(m), dim3(ML_MAX_DENSE), 0, ctx->stream, (q & 1) ? B0 : A0,
                               (q & 1) ? A0 : B0, m, q * ML_GJ_B);
    }
    hipLaunchKernelGGL(k_fcg_init, dim3(nbf), dim3(KB), 0, ctx->stream, n, S.rhs, x, rb[0], r64, x64, S.dinv32, R,
                       D.pu);
    ASD2_LAUNCH_CHECK();
    ASD2_TRY(tail<T_FCG_INIT>(&D, nbf, D.pu));
    int it = 0, fallbacks = 0;
    bool use_net = true;  // the coarse correction is in use (false after a fallback to Jacobi)
    bool fresh = true;    // r was (re)set: the next iteration starts with a standalone fine pass
    int nbu = nbf;        // blocks of the kernel whose (r.r) partials decide the next iteration
    // directions kept: the network's coarse step varies with r (gamma), so the learned
    // preconditioner keeps FCG_M; the V-cycle's is the same linear map in every iteration and
    // one direction makes the recurrence PCG
    const int mwin = (learned && !hybrid) ? FCG_M : (hybrid ? HYBRID_M : 1);
    const bool many = mwin > 1;  // the kernels' direction count: FCG_M or 1
    // the multilevel solve smooths the fine level on both sides of its coarse correction
    // (k_lp_spmv_mult); a Jacobi fallback returns to the additive kernels
    const bool mult = !learned && !many && Hh.n >= 1;
    int wstart = 0;  // first iteration of the current window of directions
    auto nwin = [&](int k) { return std::max(0, std::min(mwin, k - wstart)); };
    const Cheb ch = ml_cheb();  // the V-cycle's smoother
    const float* ycoarse = (learned && !hybrid) ? static_cast<const float*>(N->y) : yml;
    const int first = learned ? 0 : ml_first_lower(Hh);
    // recording of one level's coarse-correction input (training data), armed by
    // asd2_poisson_record_coarse for this solve
    int rec_level = -1;
    if (ctx->rec_dst) {
        const int top = learned ? (hybrid ? NL : 0) : first;
        for (int l = 0; l <= top && l < std::max(Hh.n, 1); l++) {
            const int H = Hh.n > 0 ? Hh.L[l].H : C.Hc, W = Hh.n > 0 ? Hh.L[l].W : C.Wc;
            if (H == ctx->rec_h && W == ctx->rec_w) rec_level = l;
        }
        if (rec_level < 0) {
            ctx->rec_dst = nullptr;
            ctx->rec_iters.clear();
            set_error("no level of %d x %d cells whose input this solve computes", ctx->rec_h, ctx->rec_w);
            return ASD2_ERR_ARG;
        }
    }
    struct RecordClear {
        asd2_ctx* c;
        ~RecordClear() {
            c->rec_dst = nullptr;
            c->rec_iters.clear();
        }
    } record_clear{ctx};
    auto record = [&](int level, int iter) -> int {
        if (level != rec_level) return ASD2_OK;
        const double* src = (Hh.n > 0) ? Hh.L[level].g : C.g;
        const int64_t cells = static_cast<int64_t>(ctx->rec_h) * ctx->rec_w;
        for (size_t k = 0; k < ctx->rec_iters.size(); k++)
            if (ctx->rec_iters[k] == iter)
                ASD2_HIP(hipMemcpyAsync(ctx->rec_dst + static_cast<int64_t>(k) * cells, src, cells * 8,
                                        hipMemcpyDeviceToDevice, ctx->stream));
        return ASD2_OK;
    };
    // count iterations; the last update pass leaves the state in snap for the host's poll
    auto batch = [&](int count, KrylovState* snap) -> int {
        for (int k = 0; k < count; k++, it++) {
            float* r = rb[it & 1];
            if (fresh) {
                if (mult && use_net)
                    launch_lp_pass<false, 1, true>(ctx->stream, nbb, G, C, r, nullptr, x, dirs_for(R, it), S.dinv32,
                                                   S.ax32, S.ay32, D.st, D.ps, nbs, nullptr, it, nwin(it), pb, nullptr,
                                                   ML_OMEGA, pm);
                else
                    (many ? launch_lp_pass<false, FCG_M> : launch_lp_pass<false, 1>)(ctx->stream, nbb, G, C, r, nullptr, x, dirs_for(R, it), S.dinv32, S.ax32, S.ay32,
                                      D.st, D.ps, nbs, nullptr, it, nwin(it), pb, nullptr, 0.0, nullptr);
                fresh = false;
            }
            if (use_net) {
                prof_begin(ctx, ASD2_KID_COARSE);
                ASD2_TRY(record(0, it));
                if (hybrid) {
                    for (int l = 0; l < NL; l++) {
                        if (ml_small_tiles(Hh.L[l], l)) {
                            const int nt = ceil_div(Hh.L[l].H, ML_TS) * ceil_div(Hh.L[l].W, ML_TS);
                            hipLaunchKernelGGL(k_ml_down<ML_TS>, dim3(std::min(nt, 4096)), dim3(ML_TS * ML_TS), 0,
                                               ctx->stream, Hh.L[l], Hh.L[l + 1], D.st, ch);
                        } else {
                            const int nt = ceil_div(Hh.L[l].H, ML_T) * ceil_div(Hh.L[l].W, ML_T);
                            hipLaunchKernelGGL(k_ml_down<ML_T>, dim3(std::min(nt, 4096)), dim3(KB), 0, ctx->stream,
                                               Hh.L[l], Hh.L[l + 1], D.st, ch);
                        }
                        ASD2_TRY(record(l + 1, it));
                    }
                    hipLaunchKernelGGL(k_ml_net_pack, dim3(nbn), dim3(KB), 0, ctx->stream, Hh.L[NL], N->flin, D.st);
                    ASD2_TRY(unet_forward_linear(N, 1, false));
                    hipLaunchKernelGGL(k_ml_net_dots, dim3(nbn), dim3(KB), 0, ctx->stream, Hh.L[NL], N->y, D.st,
                                       net_part);
                    hipLaunchKernelGGL(k_ml_net_apply, dim3(nbn), dim3(KB), 0, ctx->stream, Hh.L[NL], N->y, D.st,
                                       net_part, nbn);
                    for (int l = NL - 1; l >= 0; l--) {
                        if (ml_small_tiles(Hh.L[l], l)) {
                            const int nt = ceil_div(Hh.L[l].H, ML_TS) * ceil_div(Hh.L[l].W, ML_TS);
                            hipLaunchKernelGGL(k_ml_up<ML_TS>, dim3(std::min(nt, 4096)), dim3(ML_TS * ML_TS), 0,
                                               ctx->stream, Hh.L[l], Hh.L[l + 1], static_cast<float*>(nullptr), C.rc,
                                               C.qc[0], D.st, static_cast<double*>(nullptr), ch);
                        } else {
                            hipLaunchKernelGGL(k_ml_up<ML_T>,
                                               dim3(l == 0 ? nbc : flat_blocks(static_cast<int64_t>(Hh.L[l].H) * Hh.L[l].W)),
                                               dim3(KB), 0, ctx->stream, Hh.L[l], Hh.L[l + 1],
                                               l == 0 ? yml : static_cast<float*>(nullptr), C.rc, C.qc[0], D.st,
                                               l == 0 ? pc : static_cast<double*>(nullptr), ch);
                        }
                    }
                } else if (learned) {
                    hipLaunchKernelGGL(k_lp_pack, dim3(nbc), dim3(KB), 0, ctx->stream, C, N->flin, D.st, pb, nbb);
                    ASD2_TRY(unet_forward_linear(N, 1, false));
                } else if (Hh.n == 1) {
                    hipLaunchKernelGGL(k_ml_dense_only, dim3(1), dim3(KB), 0, ctx->stream, Hh.L[0], Minv, yml, D.st);
                } else {
                    for (int l = 0; l < first; l++) {
                        if (ml_small_tiles(Hh.L[l], l)) {
                            const int nt = ceil_div(Hh.L[l].H, ML_TS) * ceil_div(Hh.L[l].W, ML_TS);
                            hipLaunchKernelGGL(k_ml_down<ML_TS>, dim3(std::min(nt, 4096)), dim3(ML_TS * ML_TS), 0,
                                               ctx->stream, Hh.L[l], Hh.L[l + 1], D.st, ch);
                        } else {
                            const int nt = ceil_div(Hh.L[l].H, ML_T) * ceil_div(Hh.L[l].W, ML_T);
                            hipLaunchKernelGGL(k_ml_down<ML_T>, dim3(std::min(nt, 4096)), dim3(KB), 0, ctx->stream,
                                               Hh.L[l], Hh.L[l + 1], D.st, ch);
                        }
                        ASD2_TRY(record(l + 1, it));
                    }
                    if (first == Hh.n - 1) {
                        hipLaunchKernelGGL(k_ml_dense_apply,
                                           dim3(ceil_div(Hh.L[first].H * Hh.L[first].W, KB / 64) + 1), dim3(KB), 0,
                                           ctx->stream, Hh.L[first], Minv, D.st,
                                           FineRows{D.pu, nbu, pb, nbb, many ? nullptr : fine_totals,
                                                    (mult && use_net) ? pm : nullptr});
                    } else {
                        hipLaunchKernelGGL(k_ml_lower, dim3(1), dim3(1024), 0, ctx->stream, Hh, first, Minv, D.st, ch);
                    }
                    for (int l = first - 1; l >= 0; l--) {
                        if (ml_small_tiles(Hh.L[l], l)) {
                            const int nt = ceil_div(Hh.L[l].H, ML_TS) * ceil_div(Hh.L[l].W, ML_TS);
                            hipLaunchKernelGGL(k_ml_up<ML_TS>, dim3(std::min(nt, 4096)), dim3(ML_TS * ML_TS), 0,
                                               ctx->stream, Hh.L[l], Hh.L[l + 1], static_cast<float*>(nullptr), C.rc,
                                               C.qc[0], D.st, static_cast<double*>(nullptr), ch);
                        } else {
                            hipLaunchKernelGGL(k_ml_up<ML_T>,
                                               dim3(l == 0 ? nbc : flat_blocks(static_cast<int64_t>(Hh.L[l].H) * Hh.L[l].W)),
                                               dim3(KB), 0, ctx->stream, Hh.L[l], Hh.L[l + 1],
                                               l == 0 ? yml : static_cast<float*>(nullptr), C.rc, C.qc[0], D.st,
                                               l == 0 ? pc : static_cast<double*>(nullptr), ch,
                                               (l == 0 && mult && use_net) ? C.aq : static_cast<const double*>(nullptr),
                                               (l == 0 && mult && use_net) ? pcm : static_cast<double*>(nullptr));
                        }
                    }
                }
                if (learned || Hh.n == 1)
                    hipLaunchKernelGGL(k_lp_coarse_dots, dim3(nbc), dim3(KB), 0, ctx->stream, C, ycoarse, D.st,
                                       nwin(it), pc, (mult && use_net) ? pcm : nullptr);
                prof_end(ctx, ASD2_KID_COARSE);
            }
            prof_begin(ctx, ASD2_KID_FCG_SPMV);
            const double* ft = (use_net && !learned && Hh.n > 1 && first == Hh.n - 1 && !many) ? fine_totals : nullptr;
            if (mult && use_net)
                launch_lp_spmv_mult(ctx->stream, nbs, G, C, S.ax32, S.ay32, S.diag32, ctx->dres_ring_only ? 1 : 0, r,
                                    S.dinv32, ycoarse, dirs_for(R, it), D.st, pb, pm, nbb, pc, pcm, nbc, D.ps, it,
                                    nwin(it), ML_OMEGA, D.pu, nbu, ft);
            else
                (many ? launch_lp_spmv<FCG_M> : launch_lp_spmv<1>)(
                    ctx->stream, nbs, G, C, S.ax32, S.ay32, S.diag32, ctx->dres_ring_only ? 1 : 0, r, S.dinv32,
                    use_net ? ycoarse : nullptr,
                    dirs_for(R, it), D.st, pb, nbb, pc, nbc, D.ps, it, nwin(it), (learned && !hybrid) ? 0.0 : 1.0, D.pu,
                    nbu, ft);
            prof_end(ctx, ASD2_KID_FCG_SPMV);
            prof_begin(ctx, ASD2_KID_FCG_UPDATE);
            if (mult && use_net)
                launch_lp_pass<true, 1, true>(ctx->stream, nbb, G, C, r, rb[(it + 1) & 1], x, dirs_for(R, it + 1),
                                              S.dinv32, S.ax32, S.ay32, D.st, D.ps, nbs, D.pu, it + 1, nwin(it + 1), pb,
                                              k + 1 == count ? snap : nullptr, ML_OMEGA, pm);
            else
                (many ? launch_lp_pass<true, FCG_M> : launch_lp_pass<true, 1>)(ctx->stream, nbb, G, C, r, rb[(it + 1) & 1], x, dirs_for(R, it + 1), S.dinv32, S.ax32, S.ay32,
                                     D.st, D.ps, nbs, D.pu, it + 1, nwin(it + 1), pb, k + 1 == count ? snap : nullptr,
                                     0.0, nullptr);
            prof_end(ctx, ASD2_KID_FCG_UPDATE);
            nbu = nbb;
        }
        ASD2_LAUNCH_CHECK();
        return ASD2_OK;
    };
    // reliable update at the paused iteration k: fold x32 into x64, r64 = b - A x64, r32 = r64,
    // and re-run iteration k keeping the directions (the A-orthogonalisation of z against them
    // does not depend on the residual history); after a switch of preconditioner the window
    // restarts instead
    auto reliable = [&](int k, bool restart) -> int {
        hipLaunchKernelGGL(k_mx_accum, dim3(nbf), dim3(KB), 0, ctx->stream, n, x64, x, D.st, 0);
        hipLaunchKernelGGL(k_true_residual, dim3(nbi), dim3(KB), 0, ctx->stream, G, S.ax, S.ay, S.diag, S.rhs, x64, r64,
                           D.st, D.ps, 0);
        ASD2_TRY(tail<T_TRUE_RES>(&D, nbi, D.ps));
        hipLaunchKernelGGL(k_fcg_resume, dim3(nbf), dim3(KB), 0, ctx->stream, n, r64, rb[k & 1], S.dinv32, D.st, D.pu);
        ASD2_TRY(tail<T_MX_RESUME>(&D, 0, D.pu));
        ASD2_LAUNCH_CHECK();
        it = k;
        fresh = true;
        nbu = nbf;
        if (restart) wstart = k;
        return ASD2_OK;
    };
    const int check_every = o->check_every > 0 ? o->check_every : 4;
    // stall: ||r||^2 has not halved within this many iterations (CG residuals plateau for a
    // while even when the method works, so the window is generous)
    const int stall_window = 400;
    double best_rr = -1.0;
    int it_best = 0;
    int slot = 0;
    bool stale[2] = {false, false};  // snapshot of a batch enqueued before a host intervention
    auto enqueue = [&](int sl) -> int {
        ASD2_TRY(batch(check_every, &D.host[sl]));  // the batch's last kernel writes the snapshot
        ASD2_HIP(hipEventRecord(D.ev[sl], ctx->stream));
        return ASD2_OK;
    };
    ASD2_TRY(enqueue(0));
    ASD2_TRY(enqueue(1));
    for (int guard = 0; guard < 100000000; guard++) {
        ASD2_HIP(hipEventSynchronize(D.ev[slot]));
        const KrylovState s = D.host[slot];
        if (stale[slot]) {
            stale[slot] = false;
        } else if (s.done == DONE_BREAKDOWN && use_net) {
            // the network direction failed: continue from the current iterate with Jacobi
            use_net = false;
            fallbacks++;
            hipLaunchKernelGGL(k_set_done, dim3(1), dim3(64), 0, ctx->stream, D.st, DONE_BREAKDOWN, DONE_PAUSED);
            ASD2_TRY(reliable(s.iter, true));
            stale[slot ^ 1] = true;
        } else if (s.done == DONE_PAUSED) {
            ASD2_TRY(reliable(s.iter, false));
            stale[slot ^ 1] = true;
        } else if (s.done != DONE_RUNNING) {
            break;
        } else if (use_net) {
            if (best_rr < 0.0 || s.rr < 0.5 * best_rr) {
                best_rr = s.rr;
                it_best = s.iter;
            } else if (s.iter - it_best > stall_window) {
                use_net = false;  // stalled: the remaining iterations use Jacobi, restarting the recurrence
                fallbacks++;
                wstart = it;
            }
        }
        ASD2_TRY(enqueue(slot));
        slot ^= 1;
    }
    KrylovState fs;
    ASD2_TRY(read_state(&D, &fs));
    hipLaunchKernelGGL(k_mx_accum, dim3(nbf), dim3(KB), 0, ctx->stream, n, x64, x, D.st, 1);
    hipLaunchKernelGGL(k_copy_interior, dim3(nbi), dim3(KB), 0, ctx->stream, G, x64, d_ps);
    ASD2_LAUNCH_CHECK();
    const int st = status_of(fs);
    ASD2_TRY(report_true_residual(ctx, &D, S, x64, &fs));
    fill_stats(fs, stats);
    if (stats) stats->fallbacks = fallbacks;
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
    if (opts->precond == ASD2_PRECOND_LEARNED || opts->precond == ASD2_PRECOND_MULTILEVEL) {
        if (opts->precision != ASD2_PREC_MIXED) {
            set_error("the learned and multilevel preconditioners run in the mixed-precision solver");
            return ASD2_ERR_UNSUPPORTED;
        }
        return coarse_fcg_solve(ctx, opts, d_ps, stats, opts->precond == ASD2_PRECOND_LEARNED);
    }
    if (opts->precision == ASD2_PREC_MIXED) return solve_mixed(ctx, opts, d_ps, stats);
    return solve_fp64(ctx, opts, d_ps, stats);
}

int asd2_set_preconditioner(asd2_ctx* ctx, asd2_unet* net) {
    ASD2_TRY(check_ctx(ctx));
    ctx->net = net;
    return ASD2_OK;
}

int asd2_poisson_record_coarse(asd2_ctx* ctx, int level_h, int level_w, int count, const int* iters, double* d_out) {
    ASD2_TRY(check_ctx(ctx));
    if (count < 0 || (count > 0 && (!iters || !d_out)) || level_h <= 0 || level_w <= 0) {
        set_error("bad argument");
        return ASD2_ERR_ARG;
    }
    ctx->rec_h = level_h;
    ctx->rec_w = level_w;
    ctx->rec_iters.assign(iters, iters + count);
    ctx->rec_dst = count > 0 ? d_out : nullptr;
    return ASD2_OK;
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

namespace asd2 {

// ------------------------------------------------------------------ Gauss-Seidel (CSR)
// PoissonSolver2D::GS_2FUniform_2D (src/poisson2d.cpp:7-88).  A row-order Gauss-Seidel sweep
// reads x[c] of this sweep for c < r and of the previous sweep for c > r.  Rows are taken level
// by level, level(r) = 1 + max level(c) over c < r in row r, so every c < r a row reads was
// finished in an earlier level; x_old keeps the previous sweep for c > r.  The sweep therefore
// gives the sequential sweep's values bit for bit.  One workgroup runs the sweeps (levels
// separated by barriers), the relaxation x = x_old + 1.2 (x - x_old) and the relative change
// (block sum, fixed order), until the change is <= 1e-6 or `sweeps` sweeps (the host bounds the
// work of one launch and launches again while the change is above 1e-6; x equals x_old after
// every relaxation, so a launch can start where the last one stopped).  out[0] = sweeps run,
// out[1] = the last relative change.
__global__ __launch_bounds__(KB) void k_csr_gs(int n, const int* __restrict__ rp, const int* __restrict__ cols,
                                               const double* __restrict__ vals, const double* __restrict__ b,
                                               double* __restrict__ x, double* __restrict__ xo,
                                               const int* __restrict__ perm, const int* __restrict__ lstart,
                                               int nlev, int sweeps, double* __restrict__ out) {
#pragma clang fp contract(off)
    const double eps = 1.0e-6, omega = 1.2;
    __shared__ double s_err;
    for (int i = threadIdx.x; i < n; i += KB) xo[i] = x[i];
    __syncthreads();
    double err = 1.0;
    int it = 0;
    while (err > eps && it < sweeps) {
        for (int l = 0; l < nlev; l++) {
            for (int q = lstart[l] + static_cast<int>(threadIdx.x); q < lstart[l + 1]; q += KB) {
                const int r = perm[q];
                double t = b[r], diag = 0.0;
                for (int p = rp[r]; p < rp[r + 1]; p++) {
                    const int c = cols[p];
                    if (c == r) diag = vals[p];
                    else t -= vals[p] * (c < r ? x[c] : xo[c]);
                }
                x[r] = t / diag;
            }
            __syncthreads();
        }
        double v[2] = {0.0, 0.0};
        for (int i = threadIdx.x; i < n; i += KB) {
            const double o = xo[i];
            const double y = o + omega * (x[i] - o);
            v[0] += (y - o) * (y - o);
            v[1] += o * o;
            x[i] = y;
            xo[i] = y;
        }
        block_sum(v);
        if (threadIdx.x == 0) s_err = sqrt(v[0] / (v[1] + 1.0e-100));
        __syncthreads();
        err = s_err;
        it++;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = static_cast<double>(it);
        out[1] = err;
    }
}

}  // namespace asd2

extern "C" {

int asd2_csr_gs(asd2_ctx* ctx, int n, const int* d_rowptr, const int* d_cols, const double* d_vals,
                const double* d_b, double* d_x, int max_iter, asd2_solve_stats* stats) {
    ASD2_TRY(check_ctx(ctx));
    if (n <= 0 || !d_rowptr || !d_cols || !d_vals || !d_b || !d_x || max_iter < 0) {
        set_error("bad CSR argument");
        return ASD2_ERR_ARG;
    }
    // level schedule of the row-order dependencies, built on the host from the structure
    std::vector<int> rp(static_cast<size_t>(n) + 1);
    ASD2_HIP(hipMemcpyAsync(rp.data(), d_rowptr, rp.size() * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ASD2_HIP(hipStreamSynchronize(ctx->stream));
    const int nnz = rp[static_cast<size_t>(n)];
    if (rp[0] != 0 || nnz < 0) {
        set_error("bad CSR row pointers");
        return ASD2_ERR_ARG;
    }
    std::vector<int> cl(static_cast<size_t>(nnz > 0 ? nnz : 1));
    if (nnz > 0)
        ASD2_HIP(hipMemcpyAsync(cl.data(), d_cols, static_cast<size_t>(nnz) * sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream));
    ASD2_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<int> level(static_cast<size_t>(n), 0);
    int nlev = 0;
    for (int r = 0; r < n; r++) {
        if (rp[r + 1] < rp[r]) {
            set_error("bad CSR row pointers");
            return ASD2_ERR_ARG;
        }
        int lv = 0;
        for (int p = rp[r]; p < rp[r + 1]; p++) {
            const int c = cl[p];
            if (c < 0 || c >= n) {
                set_error("CSR column %d out of range", c);
                return ASD2_ERR_ARG;
            }
            if (c < r) lv = std::max(lv, level[c] + 1);
        }
        level[r] = lv;
        nlev = std::max(nlev, lv + 1);
    }
    std::vector<int> lstart(static_cast<size_t>(nlev) + 1, 0), perm(static_cast<size_t>(n));
    for (int r = 0; r < n; r++) lstart[level[r] + 1]++;
    for (int l = 0; l < nlev; l++) lstart[l + 1] += lstart[l];
    {
        std::vector<int> fill(lstart.begin(), lstart.end() - 1);
        for (int r = 0; r < n; r++) perm[fill[level[r]]++] = r;
    }
    int *d_perm, *d_lstart;
    double *xo, *out;
    ASD2_TRY(ws(ctx, "gs_perm", static_cast<size_t>(n), &d_perm));
    ASD2_TRY(ws(ctx, "gs_lstart", static_cast<size_t>(nlev) + 1, &d_lstart));
    ASD2_TRY(ws(ctx, "gs_xold", static_cast<size_t>(n), &xo));
    ASD2_TRY(ws(ctx, "gs_out", 2, &out));
    ASD2_HIP(hipMemcpyAsync(d_perm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    ASD2_HIP(hipMemcpyAsync(d_lstart, lstart.data(), lstart.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    // sweeps per launch: about 2^27 row entries and barrier slots of work
    const double per_sweep = static_cast<double>(nnz) + 256.0 * nlev + n;
    const int chunk = static_cast<int>(std::max(1.0, std::min(1.0e9, 134217728.0 / per_sweep)));
    int total = 0;
    double h[2] = {0.0, 1.0};
    while (h[1] > 1.0e-6 && total < max_iter) {
        const int k = std::min(chunk, max_iter - total);
        hipLaunchKernelGGL(k_csr_gs, dim3(1), dim3(KB), 0, ctx->stream, n, d_rowptr, d_cols, d_vals, d_b, d_x, xo,
                           d_perm, d_lstart, nlev, k, out);
        ASD2_LAUNCH_CHECK();
        ASD2_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        ASD2_HIP(hipStreamSynchronize(ctx->stream));
        total += static_cast<int>(h[0]);
        if (!(h[1] == h[1])) break;  // a NaN change (zero diagonal) ends the run as in the sweep loop
    }
    if (stats) {
        stats->iterations = total;
        stats->reliable_updates = 0;
        stats->fallbacks = 0;
        stats->converged = h[1] <= 1.0e-6 ? 1 : 0;
        stats->rel_residual = h[1];
        stats->true_rel_residual = -1.0;
    }
    return ASD2_OK;
}

}  // extern "C"
