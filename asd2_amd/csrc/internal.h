// internal.h -- shared definitions of the gfx950 two-phase library (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "asd2_hip.h"

namespace asd2 {

// ------------------------------------------------------------------ errors
void set_error(const char* fmt, ...);

#define ASD2_HIP(expr)                                                                  \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::asd2::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                              __FILE__, __LINE__);                                      \
            return ASD2_ERR_HIP;                                                        \
        }                                                                               \
    } while (0)

#define ASD2_TRY(expr)                  \
    do {                                \
        int s_ = (expr);                \
        if (s_ != ASD2_OK) return s_;   \
    } while (0)

#define ASD2_LAUNCH_CHECK() ASD2_HIP(hipGetLastError())

// ------------------------------------------------------------------ geometry
// Padded, j-fastest layout of the reference: idx(i, j) = j + (ny + 2g) * i.
struct Geom {
    int nx, ny, g, sy;  // sy = ny + 2g, the stride of i
    double dx, dy;
    __host__ __device__ int64_t idx(int i, int j) const {
        return static_cast<int64_t>(j) + static_cast<int64_t>(sy) * i;
    }
    __host__ __device__ int64_t size() const {
        return static_cast<int64_t>(nx + 2 * g) * static_cast<int64_t>(sy);
    }
};

// Ghost rule of one side of the P grid, for kernels that write ghosts themselves:
// 0 = mirror copy (Neumann family), 1 = -mirror + 2c (Dirichlet family), 2 = leave.
struct GhostRule {
    int kind[4];  // W, E, S, N
    double two_c[4];
};

// ------------------------------------------------------------------ device buffers
struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace asd2

// Opaque context behind the C ABI.
struct asd2_ctx {
    asd2::Geom geo{};
    double base_x = 0.0, base_y = 0.0;
    int bc[3][4] = {{9, 9, 9, 9}, {9, 9, 9, 9}, {9, 9, 9, 9}};  // [P,U,V][W,E,S,N]
    double bcc[3][4] = {};
    hipStream_t stream = nullptr;
    std::map<std::string, asd2::Buffer> bufs;

    // assembled Poisson system state
    bool system_ready = false;
    // the FP32 flux-form remainder is nonzero only on the boundary ring (assembled systems)
    bool dres_ring_only = false;
    // network of the learned preconditioner (not owned)
    struct asd2_unet* net = nullptr;
    // recording of the coarse correction's input on one level during the next coarse-corrected
    // solve (training data): level size, iterations, destination (count x H x W doubles)
    int rec_h = 0, rec_w = 0;
    std::vector<int> rec_iters;
    double* rec_dst = nullptr;

    // fast Poisson solver: rocBLAS handle (rocblas_handle) and the key of the cached eigenbasis
    void* blas = nullptr;
    std::string fp_key;

    // Krylov polling: pinned host copies of the device scalars and their events
    void* kr_host = nullptr;
    hipEvent_t kr_ev[2] = {nullptr, nullptr};

    // events
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int prof_kid = 0;
    std::vector<hipEvent_t> prof_events;  // pairs
    int prof_used = 0;
    int prof_launches = 0;
    double prof_ms = 0.0;
};

namespace asd2 {

// named, lazily allocated, zero-initialised workspace of at least `bytes`
int workspace(asd2_ctx* ctx, const char* name, size_t bytes, void** out);
template <class T>
int ws(asd2_ctx* ctx, const char* name, size_t count, T** out) {
    void* p = nullptr;
    int s = workspace(ctx, name, count * sizeof(T), &p);
    *out = static_cast<T*>(p);
    return s;
}
int zero_async(asd2_ctx* ctx, void* p, size_t bytes);

GhostRule ghost_rule_p(const asd2_ctx* ctx, bool* all_supported);
bool p_bc_supported(const asd2_ctx* ctx);

// per-launch profiling hooks (events around launches of the selected kernel id)
void prof_begin(asd2_ctx* ctx, int kid);
void prof_end(asd2_ctx* ctx, int kid);

int check_ctx(const asd2_ctx* ctx);

// internal launchers shared across translation units
int bc_apply_p(asd2_ctx* ctx, double* d);
int bc_apply_uv(asd2_ctx* ctx, int grid_kind, double* d);
// frees the fast Poisson solver's library handle (fastpoisson.hip)
void fast_poisson_release(asd2_ctx* ctx);

inline unsigned cdiv(int64_t a, int64_t b) { return static_cast<unsigned>((a + b - 1) / b); }

}  // namespace asd2

// ------------------------------------------------------------------ device helpers
namespace asd2 {

// std::max / std::min semantics of the reference
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ int isign(double v) { return (v > 0) - (v < 0); }

// MinMod of the reference, src/levelset2d.cpp:1036-1046
__device__ __forceinline__ double minmod(double v1, double v2) {
    const double p = v1 * v2;
    if (p > 0) return (fabs(v1) <= fabs(v2)) ? v1 : v2;
    return (p <= 0) ? 0.0 : __builtin_nan("");
}

// density of the axisymmetric solver's viscous terms and line sweeps at a segment between the
// level-set samples a and b: rho_H thetaH + rho_L (1 - thetaH), thetaH the heavy share of the
// segment (1 or 0 when both samples lie in one phase), src/mac2dAxisym.cpp:687-737, 1024-1045
__device__ __forceinline__ double rho_eff_axi(double rho_h, double rho_l, double a, double b) {
#pragma clang fp contract(off)
    double thH = 0.0;
    if (a >= 0.0 && b >= 0.0) thH = 1.0;
    else if (a <= 0.0 && b <= 0.0) thH = 0.0;
    else if (a >= 0.0 && b < 0.0) thH = fabs(a) / (fabs(a) + fabs(b));
    else if (a < 0.0 && b >= 0.0) thH = fabs(b) / (fabs(a) + fabs(b));
    return rho_h * thH + rho_l * (1.0 - thH);
}

// A TH x TW tile of a padded array held in registers between its loads and its LDS stores.
// tile_fetch issues the loads (addresses clamped into the array, no branch around a load; cells
// outside the padded array read as 0) and tile_store writes the LDS tile, so a kernel that fetches
// all of its tiles before storing any waits for one memory latency instead of one per guarded load.
template <int TH, int TW, int NT>
struct TileRegs {
    static constexpr int NE = TH * TW, NL = (NE + NT - 1) / NT;
    double v[NL];
};

template <int TH, int TW, int NT>
__device__ __forceinline__ void tile_fetch(const Geom& G, const double* __restrict__ a, TileRegs<TH, TW, NT>& t,
                                           int ib, int jb) {
    const int ipad = G.nx + 2 * G.g, jpad = G.ny + 2 * G.g;
#pragma unroll
    for (int k = 0; k < TileRegs<TH, TW, NT>::NL; k++) {
        const int e = static_cast<int>(threadIdx.x) + k * NT;
        const int r = e / TW, c = e - r * TW;
        const int i = ib + r, j = jb + c;
        const bool in = e < TH * TW && i >= 0 && i < ipad && j >= 0 && j < jpad;
        const double x = a[in ? G.idx(i, j) : 0];
        t.v[k] = in ? x : 0.0;
    }
}

template <int TH, int TW, int NT>
__device__ __forceinline__ void tile_store(const TileRegs<TH, TW, NT>& t, double (*s)[TW]) {
#pragma unroll
    for (int k = 0; k < TileRegs<TH, TW, NT>::NL; k++) {
        const int e = static_cast<int>(threadIdx.x) + k * NT;
        if ((TH * TW) % NT == 0 || e < TH * TW) {
            const int r = e / TW, c = e - r * TW;
            s[r][c] = t.v[k];
        }
    }
}

// write the ghost images of an interior value at (i, j) for the P grid
__device__ __forceinline__ void write_p_ghosts(const Geom& G, const GhostRule& R, double* out,
                                               int i, int j, double val) {
    const int g = G.g;
    if (i < 2 * g && R.kind[0] != 2) {
        const double gv = R.kind[0] == 0 ? val : (-val + R.two_c[0]);
        out[G.idx(2 * g - 1 - i, j)] = gv;
    }
    if (i >= G.nx && R.kind[1] != 2) {
        const double gv = R.kind[1] == 0 ? val : (-val + R.two_c[1]);
        out[G.idx(2 * G.nx + 2 * g - 1 - i, j)] = gv;
    }
    if (j < 2 * g && R.kind[2] != 2) {
        const double gv = R.kind[2] == 0 ? val : (-val + R.two_c[2]);
        out[G.idx(i, 2 * g - 1 - j)] = gv;
    }
    if (j >= G.ny && R.kind[3] != 2) {
        const double gv = R.kind[3] == 0 ? val : (-val + R.two_c[3]);
        out[G.idx(i, 2 * G.ny + 2 * g - 1 - j)] = gv;
    }
}

}  // namespace asd2
