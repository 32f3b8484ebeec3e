// ctx.hip -- context, device memory plumbing, errors, timing.
#include <cstdarg>
#include <cstring>

#include "internal.h"

namespace asd2 {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int check_ctx(const asd2_ctx* ctx) {
    if (!ctx) {
        set_error("null context");
        return ASD2_ERR_ARG;
    }
    return ASD2_OK;
}

int zero_async(asd2_ctx* ctx, void* p, size_t bytes);

int workspace(asd2_ctx* ctx, const char* name, size_t bytes, void** out) {
    auto it = ctx->bufs.find(name);
    if (it != ctx->bufs.end() && it->second.bytes >= bytes) {
        *out = it->second.ptr;
        return ASD2_OK;
    }
    if (it != ctx->bufs.end()) {
        ASD2_HIP(hipStreamSynchronize(ctx->stream));
        ASD2_HIP(hipFree(it->second.ptr));
        ctx->bufs.erase(it);
    }
    void* p = nullptr;
    ASD2_HIP(hipMalloc(&p, bytes));
    ctx->bufs[name] = Buffer{p, bytes};
    ASD2_TRY(zero_async(ctx, p, bytes));
    *out = p;
    return ASD2_OK;
}

// Zero fill with a plain kernel (no hipMemsetAsync, so the work is capturable).
__global__ void k_zero(uint64_t* p, size_t n64, unsigned char* tail, size_t ntail) {
    size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (; i < n64; i += stride) p[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

// dst = src over `count` doubles: the reference streaming kernel for calibrating HBM byte
// counters (one read and one write per element)
__global__ __launch_bounds__(256) void k_stream_copy(double* __restrict__ dst, const double* __restrict__ src,
                                                     int64_t count) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < count; i += static_cast<int64_t>(gridDim.x) * 256)
        dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_fill(double* __restrict__ dst, int64_t count, double value) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < count; i += static_cast<int64_t>(gridDim.x) * 256)
        dst[i] = value;
}

int zero_async(asd2_ctx* ctx, void* p, size_t bytes) {
    if (bytes == 0) return ASD2_OK;
    const size_t n64 = bytes / 8;
    const size_t ntail = bytes - n64 * 8;
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>(4096, (n64 + 255) / 256 + 1));
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, ctx->stream, static_cast<uint64_t*>(p), n64,
                       static_cast<unsigned char*>(p) + n64 * 8, ntail);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

bool p_bc_supported(const asd2_ctx* ctx) {
    for (int s = 0; s < 4; s++) {
        const int k = ctx->bc[ASD2_GRID_P][s];
        if (k == ASD2_BC_PERIODIC || k == ASD2_BC_LSFREE) return false;
    }
    return true;
}

GhostRule ghost_rule_p(const asd2_ctx* ctx, bool* all_supported) {
    GhostRule r{};
    bool ok = true;
    for (int s = 0; s < 4; s++) {
        const int k = ctx->bc[ASD2_GRID_P][s];
        r.two_c[s] = 2.0 * ctx->bcc[ASD2_GRID_P][s];
        if (k == ASD2_BC_NEUMANN || k == ASD2_BC_AXISYM || k == ASD2_BC_WALL) r.kind[s] = 0;
        else if (k == ASD2_BC_DIRICHLET || k == ASD2_BC_INLET || k == ASD2_BC_OUTLET ||
                 k == ASD2_BC_PRESSURE) r.kind[s] = 1;
        else if (k == ASD2_BC_CUSTOM) r.kind[s] = 2;
        else { r.kind[s] = 2; ok = false; }
    }
    if (all_supported) *all_supported = ok;
    return r;
}

void prof_begin(asd2_ctx* ctx, int kid) {
    if (ctx->prof_kid == 0 || ctx->prof_kid != kid) return;
    if (ctx->prof_used + 2 > static_cast<int>(ctx->prof_events.size())) {
        for (int k = 0; k < 256; k++) {
            hipEvent_t e;
            // timing only, read after a stream synchronize: no system-scope fence (the default
            // one writes back the caches at every record and inflates the bracketed kernel)
            if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return;
            ctx->prof_events.push_back(e);
        }
    }
    (void)hipEventRecord(ctx->prof_events[ctx->prof_used], ctx->stream);
}

void prof_end(asd2_ctx* ctx, int kid) {
    if (ctx->prof_kid == 0 || ctx->prof_kid != kid) return;
    if (ctx->prof_used + 2 > static_cast<int>(ctx->prof_events.size())) return;
    (void)hipEventRecord(ctx->prof_events[ctx->prof_used + 1], ctx->stream);
    ctx->prof_used += 2;
}

}  // namespace asd2

using namespace asd2;

extern "C" {

const char* asd2_last_error(void) { return g_err; }

int asd2_version(void) { return 1; }

int asd2_device_count(int* count) {
    if (!count) { set_error("null count"); return ASD2_ERR_ARG; }
    ASD2_HIP(hipGetDeviceCount(count));
    return ASD2_OK;
}

int asd2_device_name(char* buf, int len) {
    if (!buf || len <= 0) { set_error("bad buffer"); return ASD2_ERR_ARG; }
    hipDeviceProp_t prop;
    ASD2_HIP(hipGetDeviceProperties(&prop, 0));
    snprintf(buf, static_cast<size_t>(len), "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
             prop.multiProcessorCount);
    return ASD2_OK;
}

int asd2_malloc(void** d_ptr, size_t bytes) {
    if (!d_ptr) { set_error("null out pointer"); return ASD2_ERR_ARG; }
    *d_ptr = nullptr;
    if (bytes == 0) bytes = 8;
    ASD2_HIP(hipMalloc(d_ptr, bytes));
    ASD2_HIP(hipMemset(*d_ptr, 0, bytes));
    return ASD2_OK;
}

int asd2_free(void* d_ptr) {
    if (d_ptr) ASD2_HIP(hipFree(d_ptr));
    return ASD2_OK;
}

int asd2_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes) {
    ASD2_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
    return ASD2_OK;
}

int asd2_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes) {
    ASD2_HIP(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
    return ASD2_OK;
}

int asd2_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes) {
    ASD2_HIP(hipMemcpy(d_dst, d_src, bytes, hipMemcpyDeviceToDevice));
    return ASD2_OK;
}

int asd2_stream_copy(asd2_ctx* ctx, double* d_dst, const double* d_src, size_t count) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_dst || !d_src) { set_error("null array"); return ASD2_ERR_ARG; }
    const int blocks = static_cast<int>(std::min<size_t>(4096, (count + 255) / 256));
    hipLaunchKernelGGL(k_stream_copy, dim3(std::max(blocks, 1)), dim3(256), 0, ctx->stream, d_dst, d_src,
                       static_cast<int64_t>(count));
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_fill(asd2_ctx* ctx, double* d_dst, size_t count, double value) {
    ASD2_TRY(check_ctx(ctx));
    if (!d_dst && count) { set_error("null array"); return ASD2_ERR_ARG; }
    if (count == 0) return ASD2_OK;
    const int blocks = static_cast<int>(std::min<size_t>(4096, (count + 255) / 256));
    hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, ctx->stream, d_dst, static_cast<int64_t>(count), value);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_device_synchronize(void) {
    ASD2_HIP(hipDeviceSynchronize());
    return ASD2_OK;
}

int asd2_create(const asd2_grid* grid, asd2_ctx** out) {
    if (!grid || !out) { set_error("null argument"); return ASD2_ERR_ARG; }
    *out = nullptr;
    if (grid->num_bc_grid != 3) {
        set_error("num_bc_grid must be 3 (HJ-WENO5 reads 3 ghost layers), got %d", grid->num_bc_grid);
        return ASD2_ERR_ARG;
    }
    if (grid->nx < 3 || grid->ny < 3 || !(grid->dx > 0) || !(grid->dy > 0)) {
        set_error("invalid grid nx=%d ny=%d dx=%g dy=%g", grid->nx, grid->ny, grid->dx, grid->dy);
        return ASD2_ERR_ARG;
    }
    const int64_t cells = static_cast<int64_t>(grid->nx + 6) * (grid->ny + 6);
    if (cells > (int64_t(1) << 31)) {
        set_error("grid too large for 32-bit row indices");
        return ASD2_ERR_ARG;
    }
    asd2_ctx* c = new asd2_ctx();
    c->geo.nx = grid->nx;
    c->geo.ny = grid->ny;
    c->geo.g = grid->num_bc_grid;
    c->geo.sy = grid->ny + 2 * grid->num_bc_grid;
    c->geo.dx = grid->dx;
    c->geo.dy = grid->dy;
    c->base_x = grid->base_x;
    c->base_y = grid->base_y;
    if (hipStreamCreate(&c->stream) != hipSuccess || hipEventCreate(&c->t0) != hipSuccess ||
        hipEventCreate(&c->t1) != hipSuccess) {
        set_error("stream/event creation failed");
        delete c;
        return ASD2_ERR_HIP;
    }
    *out = c;
    return ASD2_OK;
}

int asd2_destroy(asd2_ctx* ctx) {
    if (!ctx) return ASD2_OK;
    (void)hipStreamSynchronize(ctx->stream);
    fast_poisson_release(ctx);
    for (auto& kv : ctx->bufs) (void)hipFree(kv.second.ptr);
    for (auto e : ctx->prof_events) (void)hipEventDestroy(e);
    if (ctx->kr_host) (void)hipHostFree(ctx->kr_host);
    for (auto e : ctx->kr_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->t0) (void)hipEventDestroy(ctx->t0);
    if (ctx->t1) (void)hipEventDestroy(ctx->t1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ASD2_OK;
}

int asd2_synchronize(asd2_ctx* ctx) {
    ASD2_TRY(check_ctx(ctx));
    ASD2_HIP(hipStreamSynchronize(ctx->stream));
    return ASD2_OK;
}

int asd2_set_bc(asd2_ctx* ctx, int grid_kind, const int kinds[4], const double consts[4]) {
    ASD2_TRY(check_ctx(ctx));
    if (grid_kind < 0 || grid_kind > 2 || !kinds) { set_error("bad grid kind"); return ASD2_ERR_ARG; }
    for (int s = 0; s < 4; s++) {
        if (kinds[s] < 0 || kinds[s] > 9) { set_error("bad BC kind %d", kinds[s]); return ASD2_ERR_ARG; }
        if (grid_kind != ASD2_GRID_P && kinds[s] == ASD2_BC_PRESSURE) {
            // bc2d.cpp:23-26 rejects a constant-pressure BC on velocity
            set_error("constant pressure BC is not allowed at velocity");
            return ASD2_ERR_ARG;
        }
        ctx->bc[grid_kind][s] = kinds[s];
        ctx->bcc[grid_kind][s] = consts ? consts[s] : 0.0;
    }
    ctx->system_ready = false;
    return ASD2_OK;
}

int asd2_timer_start(asd2_ctx* ctx) {
    ASD2_TRY(check_ctx(ctx));
    ASD2_HIP(hipEventRecord(ctx->t0, ctx->stream));
    return ASD2_OK;
}

int asd2_timer_stop(asd2_ctx* ctx, double* ms) {
    ASD2_TRY(check_ctx(ctx));
    ASD2_HIP(hipEventRecord(ctx->t1, ctx->stream));
    ASD2_HIP(hipEventSynchronize(ctx->t1));
    float f = 0.0f;
    ASD2_HIP(hipEventElapsedTime(&f, ctx->t0, ctx->t1));
    if (ms) *ms = f;
    return ASD2_OK;
}

int asd2_profile_kernel(asd2_ctx* ctx, int kernel_id) {
    ASD2_TRY(check_ctx(ctx));
    ctx->prof_kid = kernel_id;
    ctx->prof_used = 0;
    ctx->prof_launches = 0;
    ctx->prof_ms = 0.0;
    return ASD2_OK;
}

int asd2_profile_read_each(asd2_ctx* ctx, float* durations_ms, int capacity, int* launches) {
    ASD2_TRY(check_ctx(ctx));
    if (capacity < 0 || (capacity > 0 && !durations_ms)) { set_error("bad buffer"); return ASD2_ERR_ARG; }
    ASD2_HIP(hipStreamSynchronize(ctx->stream));
    int n = 0;
    for (int k = 0; k + 1 < ctx->prof_used; k += 2, n++) {
        float f = 0.0f;
        ASD2_HIP(hipEventElapsedTime(&f, ctx->prof_events[k], ctx->prof_events[k + 1]));
        if (n < capacity) durations_ms[n] = f;
    }
    ctx->prof_used = 0;
    if (launches) *launches = n;
    return ASD2_OK;
}

int asd2_profile_read(asd2_ctx* ctx, int* launches, double* total_ms) {
    ASD2_TRY(check_ctx(ctx));
    ASD2_HIP(hipStreamSynchronize(ctx->stream));
    double tot = 0.0;
    int n = 0;
    for (int k = 0; k + 1 < ctx->prof_used; k += 2) {
        float f = 0.0f;
        ASD2_HIP(hipEventElapsedTime(&f, ctx->prof_events[k], ctx->prof_events[k + 1]));
        tot += f;
        n++;
    }
    ctx->prof_used = 0;
    if (launches) *launches = n;
    if (total_ms) *total_ms = tot;
    return ASD2_OK;
}

}  // extern "C"
