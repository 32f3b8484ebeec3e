// unet.hip -- the learned preconditioner's network: forward, backward and Adam training,
// hand-written for gfx950 (BF16 MFMA 16x16x32 with FP32 accumulation).
//
// 3x3 convolutions are implicit GEMMs in transposed form, D[c_out][pixel] = W[c_out][k] X[k][pixel]
// with k = (tap, c_in) flattened: the weights are the MFMA A operand and 16 consecutive pixels
// the B operand, so each lane ends with 4 consecutive output channels of one pixel and stores
// them as one 8-byte NHWC write.  The input-gradient pass is the same kernel on mirrored,
// transposed weights; the weight gradient is a K = pixels GEMM staged through LDS.
#include <algorithm>
#include <cmath>

#include "solver.h"
#include "unet.h"

namespace asd2 {

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

UNetParamLayout unet_layout() {
    UNetParamLayout L{};
    int64_t off = 0;
    for (int l = 0; l < UNET_NCONV; l++) {
        const ConvSpec& c = UNET_CONVS[l];
        L.w[l] = off;
        off += 9LL * c.cout * (c.cin_a + c.cin_b);
    }
    for (int l = 0; l < UNET_NCONV; l++) {
        L.b[l] = off;
        off += UNET_CONVS[l].cout;
    }
    L.out_w = off;
    off += UNET_CONVS[UNET_NCONV - 1].cout;
    L.out_b = off;
    off += 1;
    L.total = off;
    return L;
}

// ------------------------------------------------------------------ BF16 packing
__device__ __forceinline__ bf8v pack8(uint2 lo, uint2 hi) {
    u4v u = {lo.x, lo.y, hi.x, hi.y};
    return __builtin_bit_cast(bf8v, u);
}
__device__ __forceinline__ uint2 pack4(float a, float b, float c, float d) {
    return make_uint2(static_cast<uint32_t>(f2bf(a)) | (static_cast<uint32_t>(f2bf(b)) << 16),
                      static_cast<uint32_t>(f2bf(c)) | (static_cast<uint32_t>(f2bf(d)) << 16));
}
__device__ __forceinline__ float lo16(uint32_t u) { return bf2f(static_cast<uint16_t>(u & 0xffffu)); }
__device__ __forceinline__ float hi16(uint32_t u) { return bf2f(static_cast<uint16_t>(u >> 16)); }

// ------------------------------------------------------------------ tensors
// Input of a convolution: [A (H x W, ca channels), nearest-up(B) (H/2 x W/2, cb channels)].
// pool = 1: A is read from the 2H x 2W tensor it is the 2 x 2 mean of, computed on the fly with the
// arithmetic of k_pool2 (same sums, same rounding), so no pooled copy is needed.
struct Src {
    const uint16_t* a;
    int ca;
    const uint16_t* b;
    int cb;
    int pool = 0;
};

// channels [c, c + 4) at pixel (img, i, j) of an H x W image stack; zeros outside the image
__device__ __forceinline__ uint2 load4(const Src& s, int img, int i, int j, int H, int W, int c) {
    if (i < 0 || i >= H || j < 0 || j >= W) return make_uint2(0u, 0u);
    if (c < s.ca) {
        if (s.pool) {
            float t[4] = {0.f, 0.f, 0.f, 0.f};
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) {
                    const int64_t p = (static_cast<int64_t>(img) * (2 * H) + 2 * i + a) * (2 * W) + 2 * j + b;
                    const uint2 v = *reinterpret_cast<const uint2*>(s.a + p * s.ca + c);
                    t[0] += lo16(v.x);
                    t[1] += hi16(v.x);
                    t[2] += lo16(v.y);
                    t[3] += hi16(v.y);
                }
            return pack4(0.25f * t[0], 0.25f * t[1], 0.25f * t[2], 0.25f * t[3]);
        }
        const int64_t p = (static_cast<int64_t>(img) * H + i) * W + j;
        return *reinterpret_cast<const uint2*>(s.a + p * s.ca + c);
    }
    const int H2 = H >> 1, W2 = W >> 1;
    const int64_t p = (static_cast<int64_t>(img) * H2 + (i >> 1)) * W2 + (j >> 1);
    return *reinterpret_cast<const uint2*>(s.b + p * s.cb + (c - s.ca));
}

__device__ __forceinline__ void pixel_of(int64_t p, int H, int W, int& img, int& i, int& j) {
    const int64_t hw = static_cast<int64_t>(H) * W;
    img = static_cast<int>(p / hw);
    const int64_t rem = p - img * hw;
    i = static_cast<int>(rem / W);
    j = static_cast<int>(rem - static_cast<int64_t>(i) * W);
}

// ------------------------------------------------------------------ 3x3 convolution
constexpr int CONV_NP = 4;  // 16-pixel tiles per wave (the weights are loaded once for all)

enum ConvEpi { EPI_RELU = 0, EPI_MASK = 1, EPI_NONE = 2 };

// out[p][co] = epi(bias[co] + sum_{tap, ci} in[p + d(tap)][ci] w[tap][co][ci]), d(tap) =
// (tap / 3 - 1, tap % 3 - 1).  EPI_RELU: forward; EPI_MASK: multiply by (mask[p][co] > 0),
// the ReLU derivative of the tensor this gradient belongs to; EPI_NONE: raw.
template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(256) void k_conv3x3(Src in, const uint16_t* __restrict__ w,
                                                 const float* __restrict__ bias, uint16_t* __restrict__ out,
                                                 const uint16_t* __restrict__ mask, int nimg, int H, int W) {
    static_assert(COUT % 16 == 0 && CIN % 4 == 0, "channel counts");
    constexpr int KTOT = 9 * CIN, KCH = (KTOT + 31) / 32, MT = COUT / 16;
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const int64_t P = static_cast<int64_t>(nimg) * H * W;
    const int64_t tile0 = (static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6)) * CONV_NP;
    if (tile0 * 16 >= P) return;
    int img[CONV_NP], pi[CONV_NP], pj[CONV_NP];
    bool ok[CONV_NP];
#pragma unroll
    for (int t = 0; t < CONV_NP; t++) {
        const int64_t p = (tile0 + t) * 16 + n;
        ok[t] = p < P;
        pixel_of(ok[t] ? p : 0, H, W, img[t], pi[t], pj[t]);
    }
    f4v acc[CONV_NP][MT];
#pragma unroll
    for (int t = 0; t < CONV_NP; t++)
#pragma unroll
        for (int m = 0; m < MT; m++) acc[t][m] = f4v{0.f, 0.f, 0.f, 0.f};

    for (int kc = 0; kc < KCH; kc++) {
        const int k0 = 32 * kc + 8 * g, k1 = k0 + 4;
        const int t0 = k0 / CIN, c0 = k0 - t0 * CIN;
        const int t1 = k1 / CIN, c1 = k1 - t1 * CIN;
        bf8v a[MT];
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const int co = 16 * m + n;
            const uint2 lo = (k0 < KTOT) ? *reinterpret_cast<const uint2*>(w + (static_cast<int64_t>(t0) * COUT + co) * CIN + c0)
                                         : make_uint2(0u, 0u);
            const uint2 hi = (k1 < KTOT) ? *reinterpret_cast<const uint2*>(w + (static_cast<int64_t>(t1) * COUT + co) * CIN + c1)
                                         : make_uint2(0u, 0u);
            a[m] = pack8(lo, hi);
        }
#pragma unroll
        for (int t = 0; t < CONV_NP; t++) {
            const uint2 lo = (k0 < KTOT && ok[t]) ? load4(in, img[t], pi[t] + t0 / 3 - 1, pj[t] + t0 % 3 - 1, H, W, c0)
                                                  : make_uint2(0u, 0u);
            const uint2 hi = (k1 < KTOT && ok[t]) ? load4(in, img[t], pi[t] + t1 / 3 - 1, pj[t] + t1 % 3 - 1, H, W, c1)
                                                  : make_uint2(0u, 0u);
            const bf8v b = pack8(lo, hi);
#pragma unroll
            for (int m = 0; m < MT; m++) acc[t][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b, acc[t][m], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < CONV_NP; t++) {
        if (!ok[t]) continue;
        const int64_t p = (tile0 + t) * 16 + n;
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const int co = 16 * m + 4 * g;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                v[i] = acc[t][m][i];
                if (EPI == EPI_RELU) v[i] = fmaxf(v[i] + bias[co + i], 0.0f);
            }
            if (EPI == EPI_MASK) {
                const uint2 mk = *reinterpret_cast<const uint2*>(mask + p * COUT + co);
                if (!(lo16(mk.x) > 0.f)) v[0] = 0.f;
                if (!(hi16(mk.x) > 0.f)) v[1] = 0.f;
                if (!(lo16(mk.y) > 0.f)) v[2] = 0.f;
                if (!(hi16(mk.y) > 0.f)) v[3] = 0.f;
            }
            *reinterpret_cast<uint2*>(out + p * COUT + co) = pack4(v[0], v[1], v[2], v[3]);
        }
    }
}

// The same convolution split along K for small images, where one wave per 64 pixels leaves the
// chip idle and each wave's K loop is a long chain of dependent loads: a block of KS waves owns one
// 16-pixel tile and one 16-channel output tile (blockIdx.x, blockIdx.y), wave w takes the K chunks
// w, w + KS, ..., and the KS partial tiles are added through LDS in wave order (deterministic).
template <int CIN>
constexpr int conv_ks() {
    return ((9 * CIN + 31) / 32) >= 24 ? 8 : (((9 * CIN + 31) / 32) >= 8 ? 4 : 2);
}

template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(64 * conv_ks<CIN>()) void k_conv3x3_sk(Src in, const uint16_t* __restrict__ w,
                                                                 const float* __restrict__ bias,
                                                                 uint16_t* __restrict__ out,
                                                                 const uint16_t* __restrict__ mask, int nimg, int H,
                                                                 int W) {
    static_assert(COUT % 16 == 0 && CIN % 4 == 0, "channel counts");
    constexpr int KTOT = 9 * CIN, KCH = (KTOT + 31) / 32, KS = conv_ks<CIN>();
    __shared__ f4v red[KS][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, n = lane & 15;
    const int64_t P = static_cast<int64_t>(nimg) * H * W;
    const int64_t p = static_cast<int64_t>(blockIdx.x) * 16 + n;
    const bool ok = p < P;
    int img, pi, pj;
    pixel_of(ok ? p : 0, H, W, img, pi, pj);
    const int co = 16 * blockIdx.y + n;
    f4v acc = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int kc = wv; kc < KCH; kc += KS) {
        const int k0 = 32 * kc + 8 * g, k1 = k0 + 4;
        const int t0 = k0 / CIN, c0 = k0 - t0 * CIN;
        const int t1 = k1 / CIN, c1 = k1 - t1 * CIN;
        const uint2 alo = (k0 < KTOT) ? *reinterpret_cast<const uint2*>(w + (static_cast<int64_t>(t0) * COUT + co) * CIN + c0)
                                      : make_uint2(0u, 0u);
        const uint2 ahi = (k1 < KTOT) ? *reinterpret_cast<const uint2*>(w + (static_cast<int64_t>(t1) * COUT + co) * CIN + c1)
                                      : make_uint2(0u, 0u);
        const uint2 blo = (k0 < KTOT && ok) ? load4(in, img, pi + t0 / 3 - 1, pj + t0 % 3 - 1, H, W, c0)
                                            : make_uint2(0u, 0u);
        const uint2 bhi = (k1 < KTOT && ok) ? load4(in, img, pi + t1 / 3 - 1, pj + t1 % 3 - 1, H, W, c1)
                                            : make_uint2(0u, 0u);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pack8(alo, ahi), pack8(blo, bhi), acc, 0, 0, 0);
    }
    red[wv][lane] = acc;
    __syncthreads();
    if (wv != 0 || !ok) return;
#pragma unroll
    for (int u = 1; u < KS; u++) acc += red[u][lane];
    const int c4 = 16 * blockIdx.y + 4 * g;
    float v[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[i] = acc[i];
        if (EPI == EPI_RELU) v[i] = fmaxf(v[i] + bias[c4 + i], 0.0f);
    }
    if (EPI == EPI_MASK) {
        const uint2 mk = *reinterpret_cast<const uint2*>(mask + p * COUT + c4);
        if (!(lo16(mk.x) > 0.f)) v[0] = 0.f;
        if (!(hi16(mk.x) > 0.f)) v[1] = 0.f;
        if (!(lo16(mk.y) > 0.f)) v[2] = 0.f;
        if (!(hi16(mk.y) > 0.f)) v[3] = 0.f;
    }
    *reinterpret_cast<uint2*>(out + p * COUT + c4) = pack4(v[0], v[1], v[2], v[3]);
}

// dw[tap][co][ci] += sum_p dpre[p][co] in[p + d(tap)][ci] (FP32 atomics); blockIdx.y = tap.
// A 32-pixel chunk of dpre and of the shifted input is staged transposed in LDS, then each
// wave runs K = 32 MFMAs for its (co, ci) tiles.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv_wgrad(Src in, const uint16_t* __restrict__ dpre, float* __restrict__ dw,
                                                    int nimg, int H, int W, int pix_per_block) {
    constexpr int CP = (CIN + 15) / 16 * 16, MT = COUT / 16, CT = CP / 16, NT = MT * CT, NS = (NT + 3) / 4;
    constexpr int PC = 32, LS = PC + 8;  // chunk and LDS row stride (80 B rows keep 16 B alignment)
    __shared__ __attribute__((aligned(16))) uint16_t sD[COUT][LS];
    __shared__ __attribute__((aligned(16))) uint16_t sX[CP][LS];
    const int tap = blockIdx.y, di = tap / 3 - 1, dj = tap % 3 - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, n = lane & 15;
    f4v acc[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) acc[s] = f4v{0.f, 0.f, 0.f, 0.f};
    const int64_t P = static_cast<int64_t>(nimg) * H * W;
    const int64_t pbeg = static_cast<int64_t>(blockIdx.x) * pix_per_block;
    const int64_t pend = std::min<int64_t>(P, pbeg + pix_per_block);
    for (int64_t p0 = pbeg; p0 < pend; p0 += PC) {
        for (int e = threadIdx.x; e < PC * (COUT / 4); e += 256) {
            const int q = e / (COUT / 4), c4 = (e - q * (COUT / 4)) * 4;
            const int64_t p = p0 + q;
            const uint2 v = (p < pend) ? *reinterpret_cast<const uint2*>(dpre + p * COUT + c4) : make_uint2(0u, 0u);
            sD[c4 + 0][q] = static_cast<uint16_t>(v.x & 0xffffu);
            sD[c4 + 1][q] = static_cast<uint16_t>(v.x >> 16);
            sD[c4 + 2][q] = static_cast<uint16_t>(v.y & 0xffffu);
            sD[c4 + 3][q] = static_cast<uint16_t>(v.y >> 16);
        }
        for (int e = threadIdx.x; e < PC * (CP / 4); e += 256) {
            const int q = e / (CP / 4), c4 = (e - q * (CP / 4)) * 4;
            const int64_t p = p0 + q;
            uint2 v = make_uint2(0u, 0u);
            if (p < pend && c4 < CIN) {
                int im, i, j;
                pixel_of(p, H, W, im, i, j);
                v = load4(in, im, i + di, j + dj, H, W, c4);
            }
            sX[c4 + 0][q] = static_cast<uint16_t>(v.x & 0xffffu);
            sX[c4 + 1][q] = static_cast<uint16_t>(v.x >> 16);
            sX[c4 + 2][q] = static_cast<uint16_t>(v.y & 0xffffu);
            sX[c4 + 3][q] = static_cast<uint16_t>(v.y >> 16);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int t = wave + 4 * s;
            if (t < NT) {
                const int mt = t / CT, ct = t - mt * CT;
                const bf8v a = *reinterpret_cast<const bf8v*>(&sD[16 * mt + n][8 * g]);
                const bf8v b = *reinterpret_cast<const bf8v*>(&sX[16 * ct + n][8 * g]);
                acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[s], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int t = wave + 4 * s;
        if (t >= NT) continue;
        const int mt = t / CT, ct = t - mt * CT;
        const int ci = 16 * ct + n;
        if (ci >= CIN) continue;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int co = 16 * mt + 4 * g + i;
            dw[((static_cast<int64_t>(blockIdx.x) * 9 + tap) * COUT + co) * CIN + ci] = acc[s][i];
        }
    }
}

// dst[k] += sum over blocks b (in order) of part[b stride + k], k < count
__global__ __launch_bounds__(256) void k_sum_blocks(const float* __restrict__ part, int nblk, int64_t count,
                                                    int64_t stride, float* __restrict__ dst) {
    for (int64_t k = blockIdx.x * 256LL + threadIdx.x; k < count; k += static_cast<int64_t>(gridDim.x) * 256) {
        float s = 0.f;
        for (int b = 0; b < nblk; b++) s += part[static_cast<int64_t>(b) * stride + k];
        dst[k] += s;
    }
}

// part[block][co] = sum over the block's pixels of dpre[p][co]
template <int COUT>
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t* __restrict__ dpre, float* __restrict__ db, int64_t P) {
    constexpr int R = 256 / COUT;  // pixel lanes per channel
    const int c = threadIdx.x % COUT, r = threadIdx.x / COUT;
    float s = 0.f;
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * R + r; p < P; p += static_cast<int64_t>(gridDim.x) * R)
        s += bf2f(dpre[p * COUT + c]);
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (r == 0) {
        float t = 0.f;
        for (int k = 0; k < R; k++) t += sh[k * COUT + c];
        db[static_cast<int64_t>(blockIdx.x) * COUT + c] = t;
    }
}

// ------------------------------------------------------------------ pooling, unconcat, masks
// out (H/2 x W/2) = 2x2 mean of in (H x W), C channels
__global__ __launch_bounds__(256) void k_pool2(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int nimg,
                                               int H, int W, int C) {
    const int H2 = H / 2, W2 = W / 2, C4 = C / 4;
    const int64_t total = static_cast<int64_t>(nimg) * H2 * W2 * C4;
    for (int64_t e = blockIdx.x * 256LL + threadIdx.x; e < total; e += static_cast<int64_t>(gridDim.x) * 256) {
        const int c = static_cast<int>(e % C4) * 4;
        const int64_t q = e / C4;
        int im, i, j;
        pixel_of(q, H2, W2, im, i, j);
        float s[4] = {0.f, 0.f, 0.f, 0.f};
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                const int64_t p = (static_cast<int64_t>(im) * H + 2 * i + a) * W + 2 * j + b;
                const uint2 v = *reinterpret_cast<const uint2*>(in + p * C + c);
                s[0] += lo16(v.x);
                s[1] += hi16(v.x);
                s[2] += lo16(v.y);
                s[3] += hi16(v.y);
            }
        *reinterpret_cast<uint2*>(out + q * C + c) = pack4(0.25f * s[0], 0.25f * s[1], 0.25f * s[2], 0.25f * s[3]);
    }
}

// Gradient of the deep part of a concatenated input: dX (H x W, ca + cb channels) ->
// out (H/2 x W/2, cb) = (sum over the 4 children of dX[.][ca + c]) * (mask > 0)
__global__ __launch_bounds__(256) void k_unconcat_deep(const uint16_t* __restrict__ dX, int ca, int cb,
                                                       const uint16_t* __restrict__ mask, uint16_t* __restrict__ out,
                                                       int nimg, int H, int W) {
    const int H2 = H / 2, W2 = W / 2, C4 = cb / 4, CT = ca + cb;
    const int64_t total = static_cast<int64_t>(nimg) * H2 * W2 * C4;
    for (int64_t e = blockIdx.x * 256LL + threadIdx.x; e < total; e += static_cast<int64_t>(gridDim.x) * 256) {
        const int c = static_cast<int>(e % C4) * 4;
        const int64_t q = e / C4;
        int im, i, j;
        pixel_of(q, H2, W2, im, i, j);
        float s[4] = {0.f, 0.f, 0.f, 0.f};
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                const int64_t p = (static_cast<int64_t>(im) * H + 2 * i + a) * W + 2 * j + b;
                const uint2 v = *reinterpret_cast<const uint2*>(dX + p * CT + ca + c);
                s[0] += lo16(v.x);
                s[1] += hi16(v.x);
                s[2] += lo16(v.y);
                s[3] += hi16(v.y);
            }
        const uint2 mk = *reinterpret_cast<const uint2*>(mask + q * cb + c);
        if (!(lo16(mk.x) > 0.f)) s[0] = 0.f;
        if (!(hi16(mk.x) > 0.f)) s[1] = 0.f;
        if (!(lo16(mk.y) > 0.f)) s[2] = 0.f;
        if (!(hi16(mk.y) > 0.f)) s[3] = 0.f;
        *reinterpret_cast<uint2*>(out + q * cb + c) = pack4(s[0], s[1], s[2], s[3]);
    }
}

// Gradient into an encoder output e (H x W, C): (skip part of the decoder's concat input,
// channels [0, C) of dX with stride cs) + pool2 backward of dpool (H/2 x W/2, C), times
// (e > 0).
__global__ __launch_bounds__(256) void k_skip_pool_mask(const uint16_t* __restrict__ dX, int cs,
                                                        const uint16_t* __restrict__ dpool,
                                                        const uint16_t* __restrict__ e, uint16_t* __restrict__ out,
                                                        int nimg, int H, int W, int C) {
    const int C4 = C / 4, H2 = H / 2, W2 = W / 2;
    const int64_t total = static_cast<int64_t>(nimg) * H * W * C4;
    for (int64_t k = blockIdx.x * 256LL + threadIdx.x; k < total; k += static_cast<int64_t>(gridDim.x) * 256) {
        const int c = static_cast<int>(k % C4) * 4;
        const int64_t p = k / C4;
        int im, i, j;
        pixel_of(p, H, W, im, i, j);
        const int64_t q = (static_cast<int64_t>(im) * H2 + i / 2) * W2 + j / 2;
        const uint2 a = *reinterpret_cast<const uint2*>(dX + p * cs + c);
        const uint2 b = *reinterpret_cast<const uint2*>(dpool + q * C + c);
        const uint2 mk = *reinterpret_cast<const uint2*>(e + p * C + c);
        float s[4] = {lo16(a.x) + 0.25f * lo16(b.x), hi16(a.x) + 0.25f * hi16(b.x), lo16(a.y) + 0.25f * lo16(b.y),
                      hi16(a.y) + 0.25f * hi16(b.y)};
        if (!(lo16(mk.x) > 0.f)) s[0] = 0.f;
        if (!(hi16(mk.x) > 0.f)) s[1] = 0.f;
        if (!(lo16(mk.y) > 0.f)) s[2] = 0.f;
        if (!(hi16(mk.y) > 0.f)) s[3] = 0.f;
        *reinterpret_cast<uint2*>(out + p * C + c) = pack4(s[0], s[1], s[2], s[3]);
    }
}

// ------------------------------------------------------------------ output layer
// y[p] = b + sum_c d0[p][c] w[c]  (C = 16)
__global__ __launch_bounds__(256) void k_out_fwd(const uint16_t* __restrict__ d0, const float* __restrict__ w,
                                                 const float* __restrict__ b, float* __restrict__ y, int64_t P) {
    for (int64_t p = blockIdx.x * 256LL + threadIdx.x; p < P; p += static_cast<int64_t>(gridDim.x) * 256) {
        const uint4 v0 = *reinterpret_cast<const uint4*>(d0 + p * 16);
        const uint4 v1 = *reinterpret_cast<const uint4*>(d0 + p * 16 + 8);
        const uint32_t u[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
        float s = b ? b[0] : 0.f;
#pragma unroll
        for (int k = 0; k < 8; k++) s += lo16(u[k]) * w[2 * k] + hi16(u[k]) * w[2 * k + 1];
        y[p] = s;
    }
}

// dd0[p][c] = dy[p] w[c] (m0 > 0); part[block] = (sum_p dy d0[p][c] for c < 16, sum_p dy);
// m0 = d0 for the ReLU pass, the ReLU pass's activation for the linear pass
__global__ __launch_bounds__(256) void k_out_bwd(const uint16_t* __restrict__ d0, const uint16_t* __restrict__ m0,
                                                 const float* __restrict__ w, const float* __restrict__ dy,
                                                 uint16_t* __restrict__ dd0, float* __restrict__ gw,
                                                 float* __restrict__ gb, int64_t P) {
    float aw[16] = {}, ab = 0.f;
    for (int64_t p = blockIdx.x * 256LL + threadIdx.x; p < P; p += static_cast<int64_t>(gridDim.x) * 256) {
        const uint4 v0 = *reinterpret_cast<const uint4*>(d0 + p * 16);
        const uint4 v1 = *reinterpret_cast<const uint4*>(d0 + p * 16 + 8);
        const uint4 m0v = *reinterpret_cast<const uint4*>(m0 + p * 16);
        const uint4 m1v = *reinterpret_cast<const uint4*>(m0 + p * 16 + 8);
        const uint32_t u[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
        const uint32_t mu[8] = {m0v.x, m0v.y, m0v.z, m0v.w, m1v.x, m1v.y, m1v.z, m1v.w};
        const float g = dy[p];
        uint32_t o[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float a = lo16(u[k]), c = hi16(u[k]);
            aw[2 * k] += g * a;
            aw[2 * k + 1] += g * c;
            const float ga = lo16(mu[k]) > 0.f ? g * w[2 * k] : 0.f, gc = hi16(mu[k]) > 0.f ? g * w[2 * k + 1] : 0.f;
            o[k] = static_cast<uint32_t>(f2bf(ga)) | (static_cast<uint32_t>(f2bf(gc)) << 16);
        }
        ab += g;
        *reinterpret_cast<uint4*>(dd0 + p * 16) = make_uint4(o[0], o[1], o[2], o[3]);
        *reinterpret_cast<uint4*>(dd0 + p * 16 + 8) = make_uint4(o[4], o[5], o[6], o[7]);
    }
    __shared__ float sh[17][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 17; k++) {
        float v = k < 16 ? aw[k] : ab;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) sh[k][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < 17) {
        const float t = sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3];
        gw[static_cast<int64_t>(blockIdx.x) * 17 + threadIdx.x] = t;
    }
    (void)gb;
}

// ------------------------------------------------------------------ weights and Adam
// BF16 copies: wf = W[tap][co][ci], wb = W[8 - tap][co][ci] stored [tap][ci][co]
__global__ void k_refresh_weights(const float* __restrict__ w32, uint16_t* __restrict__ wf, uint16_t* __restrict__ wb,
                                  int64_t off, int cin, int cout) {
    const int64_t n = 9LL * cin * cout;
    for (int64_t e = blockIdx.x * 256LL + threadIdx.x; e < n; e += static_cast<int64_t>(gridDim.x) * 256) {
        const int tap = static_cast<int>(e / (static_cast<int64_t>(cin) * cout));
        const int rem = static_cast<int>(e - static_cast<int64_t>(tap) * cin * cout);
        const int co = rem / cin, ci = rem - co * cin;
        const uint16_t v = f2bf(w32[off + e]);
        wf[off + e] = v;
        wb[off + (static_cast<int64_t>(8 - tap) * cin + ci) * cout + co] = v;
    }
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2, float gscale) {
    for (int64_t k = blockIdx.x * 256LL + threadIdx.x; k < n; k += static_cast<int64_t>(gridDim.x) * 256) {
        const float gk = g[k] * gscale;
        const float mk = b1 * m[k] + (1.f - b1) * gk;
        const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
        m[k] = mk;
        v[k] = vk;
        w[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    }
}

// ------------------------------------------------------------------ training systems and energy loss
// Systems are FP32 in the padded layout of the solver ((nx + 6) (ny + 6) values per system,
// idx = (i + 3) sy + j + 3 for interior (i, j), sy = ny + 6); the network sees the coarse grid
// of c x c blocks.  Per sample b, sums[TS * b + k] holds k = 0: r.D^-1 r, 1: (D^-1 r).(r - A D^-1 r),
// 2: sum (g / dc)^2, 3: y.g, 4: y.A_c y.
constexpr int TS = 6;

// the block's sums of NV per-thread values (wave shuffles, then the four waves in order) into
// sp[((b nblk + block) TS + s0 + k)]; k_tr_finish adds the blocks in order
template <int NV>
__device__ __forceinline__ void block_sums(const float (&v)[NV], float* __restrict__ sp, int b, int s0) {
    __shared__ float sh[NV][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        float x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[k][wv] = x;
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < NV) {
        const int k = threadIdx.x;
        sp[(static_cast<int64_t>(b) * gridDim.x + blockIdx.x) * TS + s0 + k] = sh[k][0] + sh[k][1] + sh[k][2] + sh[k][3];
    }
}

// sums[TS b + s] = sum over the nblk blocks (in order) of the partials of slots s0 .. s0 + nv - 1;
// slot 0 = energy[b] instead when energy is given
__global__ __launch_bounds__(64) void k_tr_finish(const float* __restrict__ sp, int nblk, int batch, int s0, int nv,
                                                  const float* __restrict__ energy, float* __restrict__ sums) {
    for (int t = threadIdx.x; t < batch * nv; t += 64) {
        const int b = t / nv, s = s0 + t % nv;
        float v = 0.f;
        for (int k = 0; k < nblk; k++) v += sp[(static_cast<int64_t>(b) * nblk + k) * TS + s];
        if (energy && s == 0) v = energy[b];
        sums[TS * b + s] = v;
    }
}

// coarse operators of the batch (blockIdx.y = sample)
__global__ __launch_bounds__(256) void k_tr_coarse_op(const float* __restrict__ ax, const float* __restrict__ ay,
                                                      const float* __restrict__ dg, int64_t stride, int sy, int c,
                                                      int Hc, int Wc, float* __restrict__ acx, float* __restrict__ acy,
                                                      float* __restrict__ dc) {
    const int b = blockIdx.y;
    const int64_t cs = static_cast<int64_t>(Hc + 1) * (Wc + 1);
    const int64_t off = b * stride;
    for (int64_t e = blockIdx.x * 256LL + threadIdx.x; e < cs; e += static_cast<int64_t>(gridDim.x) * 256) {
        const int I = static_cast<int>(e / (Wc + 1)), J = static_cast<int>(e - static_cast<int64_t>(I) * (Wc + 1));
        coarse_cell(ax + off, ay + off, dg + off, sy, 3, c, Hc, Wc, I, J, acx + b * cs, acy + b * cs, dc + b * cs);
    }
}

// g[b] = P^T (r - A D^-1 r), one thread per coarse cell summing its c x c fine cells; block
// partials of sums 0 (r.D^-1 r) and 1 ((D^-1 r).(r - A D^-1 r))
__global__ __launch_bounds__(256) void k_tr_restrict(const float* __restrict__ r, const float* __restrict__ dinv,
                                                     const float* __restrict__ ax, const float* __restrict__ ay,
                                                     const float* __restrict__ dg, int64_t stride, int nx, int ny,
                                                     int c, float* __restrict__ g, float* __restrict__ sp) {
    const int b = blockIdx.y, sy = ny + 6, Wc = ny / c;
    const int64_t Pc = static_cast<int64_t>(nx / c) * Wc;
    const float* rb = r + b * stride;
    const float* db = dinv + b * stride;
    const float* axb = ax + b * stride;
    const float* ayb = ay + b * stride;
    const float* dgb = dg + b * stride;
    float* gb = g + b * Pc;
    float a[2] = {0.f, 0.f};
    for (int64_t p = blockIdx.x * 256LL + threadIdx.x; p < Pc; p += static_cast<int64_t>(gridDim.x) * 256) {
        const int I = static_cast<int>(p / Wc), J = static_cast<int>(p - static_cast<int64_t>(I) * Wc);
        float gs = 0.f;
        for (int u = 0; u < c; u++)
            for (int v = 0; v < c; v++) {
                const int i = I * c + u, j = J * c + v;
                const int64_t k = static_cast<int64_t>(i + 3) * sy + j + 3;
                auto zj = [&](int64_t m) { return db[m] * rb[m]; };
                const float z0 = zj(k);
                float az = dgb[k] * z0;
                if (i > 0) az += axb[k] * zj(k - sy);
                if (i + 1 < nx) az += axb[k + sy] * zj(k + sy);
                if (j > 0) az += ayb[k] * zj(k - 1);
                if (j + 1 < ny) az += ayb[k + 1] * zj(k + 1);
                const float w = rb[k] - az;
                gs += w;
                a[0] += z0 * rb[k];
                a[1] += z0 * w;
            }
        gb[p] = gs;
    }
    block_sums<2>(a, sp, b, 0);
}

// sums 2 = sum over the coarse cells of (g / dc)^2
__global__ __launch_bounds__(256) void k_tr_usum(const float* __restrict__ g, const float* __restrict__ dc, int Hc,
                                                 int Wc, float* __restrict__ sp) {
    const int b = blockIdx.y;
    const int64_t Pc = static_cast<int64_t>(Hc) * Wc, cs = static_cast<int64_t>(Hc + 1) * (Wc + 1);
    float a[1] = {0.f};
    for (int64_t p = blockIdx.x * 256LL + threadIdx.x; p < Pc; p += static_cast<int64_t>(gridDim.x) * 256) {
        const int I = static_cast<int>(p / Wc), J = static_cast<int>(p - static_cast<int64_t>(I) * Wc);
        const float u = g[b * Pc + p] / dc[b * cs + I * (Wc + 1) + J];
        a[0] += u * u;
    }
    block_sums<1>(a, sp, b, 2);
}

// f[b][p] = (0, log coarse W-face, S-face coefficient, log coarse diagonal): the ReLU pass;
// fl[b][p] = (g / dc / s, 0, 0, 0): the linear pass
__global__ __launch_bounds__(256) void k_tr_features(const float* __restrict__ g, const float* __restrict__ acx,
                                                     const float* __restrict__ acy, const float* __restrict__ dc, int Hc,
                                                     int Wc, float h2x, float h2y, const float* __restrict__ sums,
                                                     uint16_t* __restrict__ f, uint16_t* __restrict__ fl) {
    const int b = blockIdx.y;
    const int64_t Pc = static_cast<int64_t>(Hc) * Wc, cs = static_cast<int64_t>(Hc + 1) * (Wc + 1);
    float s = sqrtf(sums[TS * b + 2] / static_cast<float>(Pc));
    if (!(s > 0.f) || !isfinite(s)) s = 1.f;
    const float is = 1.f / s, h2 = fminf(h2x, h2y);
    for (int64_t p = blockIdx.x * 256LL + threadIdx.x; p < Pc; p += static_cast<int64_t>(gridDim.x) * 256) {
        const int I = static_cast<int>(p / Wc), J = static_cast<int>(p - static_cast<int64_t>(I) * Wc);
        const int64_t k = b * cs + I * (Wc + 1) + J;
        *reinterpret_cast<uint2*>(f + (b * Pc + p) * 4) =
            pack4(0.f, log_coeff(acx[k], h2x), log_coeff(acy[k], h2y), log_diag(dc[k], h2));
        *reinterpret_cast<uint2*>(fl + (b * Pc + p) * 4) = pack4(g[b * Pc + p] / dc[k] * is, 0.f, 0.f, 0.f);
    }
}

// sums 3, 4: y.g and y.A_c y
__global__ __launch_bounds__(256) void k_tr_gamma(const float* __restrict__ y, const float* __restrict__ g,
                                                  const float* __restrict__ acx, const float* __restrict__ acy,
                                                  const float* __restrict__ dc, int Hc, int Wc,
                                                  float* __restrict__ sp) {
    const int b = blockIdx.y;
    const int64_t Pc = static_cast<int64_t>(Hc) * Wc, cs = static_cast<int64_t>(Hc + 1) * (Wc + 1);
    float a[2] = {0.f, 0.f};
    for (int64_t p = blockIdx.x * 256LL + threadIdx.x; p < Pc; p += static_cast<int64_t>(gridDim.x) * 256) {
        const int I = static_cast<int>(p / Wc), J = static_cast<int>(p - static_cast<int64_t>(I) * Wc);
        const float yv = y[b * Pc + p];
        a[0] += yv * g[b * Pc + p];
        a[1] += yv * coarse_apply(acx + b * cs, acy + b * cs, dc + b * cs, y + b * Pc, Hc, Wc, I, J);
    }
    block_sums<2>(a, sp, b, 3);
}

__device__ __forceinline__ float gamma_from(float yg, float yay) {
    const float gm = yg / yay;
    return (yay > 0.f && isfinite(gm)) ? gm : 0.f;
}

// dy = -2 gamma (g - gamma A_c y) / (r.D^-1 r nb)
__global__ __launch_bounds__(256) void k_tr_dy(const float* __restrict__ y, const float* __restrict__ g,
                                               const float* __restrict__ acx, const float* __restrict__ acy,
                                               const float* __restrict__ dc, int Hc, int Wc,
                                               const float* __restrict__ sums, float* __restrict__ dy, int nb) {
    const int b = blockIdx.y;
    const int64_t Pc = static_cast<int64_t>(Hc) * Wc, cs = static_cast<int64_t>(Hc + 1) * (Wc + 1);
    const float gm = gamma_from(sums[TS * b + 3], sums[TS * b + 4]);
    float k2 = -2.f * gm / (sums[TS * b + 0] * static_cast<float>(nb));
    if (!isfinite(k2)) k2 = 0.f;  // r_b = 0: no contribution
    for (int64_t p = blockIdx.x * 256LL + threadIdx.x; p < Pc; p += static_cast<int64_t>(gridDim.x) * 256) {
        const int I = static_cast<int>(p / Wc), J = static_cast<int>(p - static_cast<int64_t>(I) * Wc);
        const float ay_ = coarse_apply(acx + b * cs, acy + b * cs, dc + b * cs, y + b * Pc, Hc, Wc, I, J);
        dy[b * Pc + p] = k2 * (g[b * Pc + p] - gm * ay_);
    }
}

// coarse training samples: block partials of sums 0 = g.D_c^-1 g (k_tr_finish replaces it by
// energy[b] = g.A_c^-1 g when given: the loss normaliser), 1 = 0, 2 = sum (g / dc)^2
__global__ __launch_bounds__(256) void k_tr_gnorm(const float* __restrict__ g, const float* __restrict__ dc, int Hc,
                                                  int Wc, float* __restrict__ sp) {
    const int b = blockIdx.y;
    const int64_t Pc = static_cast<int64_t>(Hc) * Wc, cs = static_cast<int64_t>(Hc + 1) * (Wc + 1);
    float a[3] = {0.f, 0.f, 0.f};
    for (int64_t p = blockIdx.x * 256LL + threadIdx.x; p < Pc; p += static_cast<int64_t>(gridDim.x) * 256) {
        const int I = static_cast<int>(p / Wc), J = static_cast<int>(p - static_cast<int64_t>(I) * Wc);
        const float gv = g[b * Pc + p], d = dc[b * cs + I * (Wc + 1) + J];
        a[0] += gv * gv / d;
        a[2] += (gv / d) * (gv / d);
    }
    block_sums<3>(a, sp, b, 0);
}

// FP32 features [P][4] -> feat (channel 0 zeroed) and flin (channel 0 only), BF16
__global__ void k_split_features(const float* __restrict__ in, uint16_t* __restrict__ f, uint16_t* __restrict__ fl,
                                 int64_t P) {
    for (int64_t p = blockIdx.x * 256LL + threadIdx.x; p < P; p += static_cast<int64_t>(gridDim.x) * 256) {
        const float4 v = *reinterpret_cast<const float4*>(in + p * 4);
        *reinterpret_cast<uint2*>(f + p * 4) = pack4(0.f, v.y, v.z, v.w);
        *reinterpret_cast<uint2*>(fl + p * 4) = pack4(v.x, 0.f, 0.f, 0.f);
    }
}

__global__ void k_f32_to_bf16(const float* __restrict__ in, uint16_t* __restrict__ out, int64_t n) {
    for (int64_t k = blockIdx.x * 256LL + threadIdx.x; k < n; k += static_cast<int64_t>(gridDim.x) * 256)
        out[k] = f2bf(in[k]);
}

// ------------------------------------------------------------------ host orchestration
namespace {

int blocks_for(int64_t work, int per = 256, int cap = 65535) {
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(cap, (work + per - 1) / per)));
}

// the K-split kernel when the pixel-tiled one would run fewer waves than the chip has SIMDs
template <int CIN, int COUT, int EPI>
void launch_conv(hipStream_t s, const Src& in, const uint16_t* w, const float* bias, uint16_t* out,
                 const uint16_t* mask, int nimg, int H, int W) {
    const int64_t tiles = (static_cast<int64_t>(nimg) * H * W + 15) / 16;
    const int64_t waves = (tiles + CONV_NP - 1) / CONV_NP;
    if (waves < 1024) {
        hipLaunchKernelGGL((k_conv3x3_sk<CIN, COUT, EPI>), dim3(static_cast<unsigned>(tiles), COUT / 16),
                           dim3(64 * conv_ks<CIN>()), 0, s, in, w, bias, out, mask, nimg, H, W);
        return;
    }
    const int blocks = static_cast<int>((waves + 3) / 4);
    hipLaunchKernelGGL((k_conv3x3<CIN, COUT, EPI>), dim3(blocks), dim3(256), 0, s, in, w, bias, out, mask, nimg, H, W);
}

// at most WG_BLOCKS pixel ranges per tap: the partial slab holds WG_BLOCKS x 9 x COUT x CIN sums
constexpr int WG_BLOCKS = 32, BG_BLOCKS = 2048;

void sum_blocks(hipStream_t s, const float* part, int nblk, int64_t count, float* dst, int64_t stride = -1) {
    hipLaunchKernelGGL(k_sum_blocks, dim3(blocks_for(count, 256, 4096)), dim3(256), 0, s, part, nblk, count,
                       stride < 0 ? count : stride, dst);
}

template <int CIN, int COUT>
void launch_wgrad(hipStream_t s, float* part, const Src& in, const uint16_t* dpre, float* dw, int nimg, int H, int W) {
    const int64_t P = static_cast<int64_t>(nimg) * H * W;
    const int64_t per = std::max<int64_t>(256, ((P + WG_BLOCKS - 1) / WG_BLOCKS + 31) / 32 * 32);
    const int blocks = static_cast<int>((P + per - 1) / per);
    hipLaunchKernelGGL((k_conv_wgrad<CIN, COUT>), dim3(blocks, 9), dim3(256), 0, s, in, dpre, part, nimg, H, W,
                       static_cast<int>(per));
    sum_blocks(s, part, blocks, 9LL * COUT * CIN, dw);
}

template <int COUT>
void launch_bgrad(hipStream_t s, float* part, const uint16_t* dpre, float* db, int64_t P) {
    if (P == 0) return;
    const int blocks = blocks_for(P, 256 / COUT * 64, BG_BLOCKS);
    hipLaunchKernelGGL(k_bias_grad<COUT>, dim3(blocks), dim3(256), 0, s, dpre, part, P);
    sum_blocks(s, part, blocks, COUT, db);
}

// forward conv l (template dispatch by layer)
void conv_forward(asd2_unet* N, int l, const Src& in, int nimg, int H, int W) {
    const uint16_t* w = N->wf + N->L.w[l];
    const float* b = N->w32 + N->L.b[l];
    uint16_t* o = N->act[l];
    hipStream_t s = N->stream;
    switch (l) {
        case 0: launch_conv<4, 16, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 1: launch_conv<16, 16, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 2: launch_conv<16, 32, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 3: launch_conv<32, 32, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 4: launch_conv<32, 64, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 5: launch_conv<64, 64, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 6: launch_conv<64, 128, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 7: launch_conv<128, 128, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 8: launch_conv<192, 64, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 9: launch_conv<64, 64, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 10: launch_conv<96, 32, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 11: launch_conv<32, 32, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 12: launch_conv<48, 16, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        case 13: launch_conv<16, 16, EPI_RELU>(s, in, w, b, o, nullptr, nimg, H, W); break;
        default: break;
    }
}

// linear-pass conv l: out = conv(in, wf[l]) * (act[l] > 0), no bias
void conv_forward_lin(asd2_unet* N, int l, const Src& in, int nimg, int H, int W) {
    const uint16_t* w = N->wf + N->L.w[l];
    uint16_t* o = N->lin[l];
    const uint16_t* m = N->act[l];
    hipStream_t s = N->stream;
    switch (l) {
        case 0: launch_conv<4, 16, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 1: launch_conv<16, 16, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 2: launch_conv<16, 32, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 3: launch_conv<32, 32, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 4: launch_conv<32, 64, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 5: launch_conv<64, 64, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 6: launch_conv<64, 128, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 7: launch_conv<128, 128, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 8: launch_conv<192, 64, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 9: launch_conv<64, 64, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 10: launch_conv<96, 32, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 11: launch_conv<32, 32, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 12: launch_conv<48, 16, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        case 13: launch_conv<16, 16, EPI_MASK>(s, in, w, nullptr, o, m, nimg, H, W); break;
        default: break;
    }
}

// input gradient of conv l: out (cin channels) = conv(dpre (cout), wb[l]) with an epilogue
void conv_dgrad(asd2_unet* N, int l, const uint16_t* dpre, uint16_t* out, const uint16_t* mask, int nimg, int H,
                int W) {
    const uint16_t* w = N->wb + N->L.w[l];
    const Src in{dpre, UNET_CONVS[l].cout, nullptr, 0};
    hipStream_t s = N->stream;
    const bool m = mask != nullptr;
    switch (l) {
        case 1: m ? launch_conv<16, 16, EPI_MASK>(s, in, w, nullptr, out, mask, nimg, H, W)
                  : launch_conv<16, 16, EPI_NONE>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 2: launch_conv<32, 16, EPI_NONE>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 3: launch_conv<32, 32, EPI_MASK>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 4: launch_conv<64, 32, EPI_NONE>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 5: launch_conv<64, 64, EPI_MASK>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 6: launch_conv<128, 64, EPI_NONE>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 7: launch_conv<128, 128, EPI_MASK>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 8: launch_conv<64, 192, EPI_NONE>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 9: launch_conv<64, 64, EPI_MASK>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 10: launch_conv<32, 96, EPI_NONE>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 11: launch_conv<32, 32, EPI_MASK>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 12: launch_conv<16, 48, EPI_NONE>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        case 13: launch_conv<16, 16, EPI_MASK>(s, in, w, nullptr, out, mask, nimg, H, W); break;
        default: break;
    }
}

void conv_wgrad(asd2_unet* N, int l, const Src& in, const uint16_t* dpre, int nimg, int H, int W, bool bias) {
    float* part = N->gpart;
    float* dw = N->grad + N->L.w[l];
    float* db = N->grad + N->L.b[l];
    hipStream_t s = N->stream;
    const int64_t P = bias ? static_cast<int64_t>(nimg) * H * W : 0;  // P = 0: no bias gradient
    switch (l) {
        case 0: launch_wgrad<4, 16>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<16>(s, part, dpre, db, P); break;
        case 1: launch_wgrad<16, 16>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<16>(s, part, dpre, db, P); break;
        case 2: launch_wgrad<16, 32>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<32>(s, part, dpre, db, P); break;
        case 3: launch_wgrad<32, 32>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<32>(s, part, dpre, db, P); break;
        case 4: launch_wgrad<32, 64>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<64>(s, part, dpre, db, P); break;
        case 5: launch_wgrad<64, 64>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<64>(s, part, dpre, db, P); break;
        case 6: launch_wgrad<64, 128>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<128>(s, part, dpre, db, P); break;
        case 7: launch_wgrad<128, 128>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<128>(s, part, dpre, db, P); break;
        case 8: launch_wgrad<192, 64>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<64>(s, part, dpre, db, P); break;
        case 9: launch_wgrad<64, 64>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<64>(s, part, dpre, db, P); break;
        case 10: launch_wgrad<96, 32>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<32>(s, part, dpre, db, P); break;
        case 11: launch_wgrad<32, 32>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<32>(s, part, dpre, db, P); break;
        case 12: launch_wgrad<48, 16>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<16>(s, part, dpre, db, P); break;
        case 13: launch_wgrad<16, 16>(s, part, in, dpre, dw, nimg, H, W); launch_bgrad<16>(s, part, dpre, db, P); break;
        default: break;
    }
}

int check_net(const asd2_unet* N) {
    if (!N) {
        set_error("null network");
        return ASD2_ERR_ARG;
    }
    return ASD2_OK;
}

// floats of the partial slab: the largest weight gradient (WG_BLOCKS x 9 x COUT x CIN), the bias
// and output-layer gradients (BG_BLOCKS x 128, 2048 x 17), the training sums (batch x 1024 x TS)
int64_t gpart_size(int max_batch) {
    int64_t n = static_cast<int64_t>(BG_BLOCKS) * 128;
    for (int l = 0; l < UNET_NCONV; l++)
        n = std::max<int64_t>(n, static_cast<int64_t>(WG_BLOCKS) * 9 * UNET_CONVS[l].cout *
                                     (UNET_CONVS[l].cin_a + UNET_CONVS[l].cin_b));
    return std::max<int64_t>(n, static_cast<int64_t>(max_batch) * 1024 * TS);
}

template <typename T>
int dalloc(T** p, int64_t count) {
    ASD2_HIP(hipMalloc(reinterpret_cast<void**>(p), std::max<int64_t>(count, 1) * sizeof(T)));
    ASD2_HIP(hipMemset(*p, 0, std::max<int64_t>(count, 1) * sizeof(T)));
    return ASD2_OK;
}

// gradient scratch: [0..19] as listed in unet_backward
struct GradBufs {
    uint16_t* t[20];
};

}  // namespace

int unet_refresh(asd2_unet* N) {
    for (int l = 0; l < UNET_NCONV; l++) {
        const ConvSpec& c = UNET_CONVS[l];
        const int cin = c.cin_a + c.cin_b;
        hipLaunchKernelGGL(k_refresh_weights, dim3(blocks_for(9LL * cin * c.cout, 256, 1024)), dim3(256), 0, N->stream,
                           N->w32, N->wf, N->wb, N->L.w[l], cin, c.cout);
    }
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

// forward of nimg images from the feature tensor N->feat; y = N->y
int unet_forward(asd2_unet* N, int nimg) {
    const int H = N->H, W = N->W;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    hipStream_t s = N->stream;
    uint16_t** a = N->act;
    conv_forward(N, 0, Src{N->feat, 4, nullptr, 0}, nimg, H, W);
    conv_forward(N, 1, Src{a[0], 16, nullptr, 0}, nimg, H, W);
    hipLaunchKernelGGL(k_pool2, dim3(blocks_for(static_cast<int64_t>(nimg) * H1 * W1 * 4)), dim3(256), 0, s, a[1],
                       N->pooled[0], nimg, H, W, 16);
    conv_forward(N, 2, Src{N->pooled[0], 16, nullptr, 0}, nimg, H1, W1);
    conv_forward(N, 3, Src{a[2], 32, nullptr, 0}, nimg, H1, W1);
    hipLaunchKernelGGL(k_pool2, dim3(blocks_for(static_cast<int64_t>(nimg) * H2 * W2 * 8)), dim3(256), 0, s, a[3],
                       N->pooled[1], nimg, H1, W1, 32);
    conv_forward(N, 4, Src{N->pooled[1], 32, nullptr, 0}, nimg, H2, W2);
    conv_forward(N, 5, Src{a[4], 64, nullptr, 0}, nimg, H2, W2);
    hipLaunchKernelGGL(k_pool2, dim3(blocks_for(static_cast<int64_t>(nimg) * H3 * W3 * 16)), dim3(256), 0, s, a[5],
                       N->pooled[2], nimg, H2, W2, 64);
    conv_forward(N, 6, Src{N->pooled[2], 64, nullptr, 0}, nimg, H3, W3);
    conv_forward(N, 7, Src{a[6], 128, nullptr, 0}, nimg, H3, W3);
    conv_forward(N, 8, Src{a[5], 64, a[7], 128}, nimg, H2, W2);
    conv_forward(N, 9, Src{a[8], 64, nullptr, 0}, nimg, H2, W2);
    conv_forward(N, 10, Src{a[3], 32, a[9], 64}, nimg, H1, W1);
    conv_forward(N, 11, Src{a[10], 32, nullptr, 0}, nimg, H1, W1);
    conv_forward(N, 12, Src{a[1], 16, a[11], 32}, nimg, H, W);
    conv_forward(N, 13, Src{a[12], 16, nullptr, 0}, nimg, H, W);
    const int64_t P = static_cast<int64_t>(nimg) * H * W;
    hipLaunchKernelGGL(k_out_fwd, dim3(blocks_for(P, 256, 8192)), dim3(256), 0, s, a[13], N->w32 + N->L.out_w,
                       N->w32 + N->L.out_b, N->y, P);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int unet_forward_linear(asd2_unet* N, int nimg, bool keep_pooled) {
    const int H = N->H, W = N->W;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    hipStream_t s = N->stream;
    uint16_t** a = N->lin;
    // the pooled inputs of convs 2, 4, 6: stored (the backward pass reads them) or formed in the
    // convolution's loads
    auto pooled = [&](int k, const uint16_t* fine, int c, int Hf, int Wf) -> Src {
        if (!keep_pooled) return Src{fine, c, nullptr, 0, 1};
        hipLaunchKernelGGL(k_pool2, dim3(blocks_for(static_cast<int64_t>(nimg) * (Hf / 2) * (Wf / 2) * (c / 4))),
                           dim3(256), 0, s, fine, N->linpooled[k], nimg, Hf, Wf, c);
        return Src{N->linpooled[k], c, nullptr, 0};
    };
    conv_forward_lin(N, 0, Src{N->flin, 4, nullptr, 0}, nimg, H, W);
    conv_forward_lin(N, 1, Src{a[0], 16, nullptr, 0}, nimg, H, W);
    conv_forward_lin(N, 2, pooled(0, a[1], 16, H, W), nimg, H1, W1);
    conv_forward_lin(N, 3, Src{a[2], 32, nullptr, 0}, nimg, H1, W1);
    conv_forward_lin(N, 4, pooled(1, a[3], 32, H1, W1), nimg, H2, W2);
    conv_forward_lin(N, 5, Src{a[4], 64, nullptr, 0}, nimg, H2, W2);
    conv_forward_lin(N, 6, pooled(2, a[5], 64, H2, W2), nimg, H3, W3);
    conv_forward_lin(N, 7, Src{a[6], 128, nullptr, 0}, nimg, H3, W3);
    conv_forward_lin(N, 8, Src{a[5], 64, a[7], 128}, nimg, H2, W2);
    conv_forward_lin(N, 9, Src{a[8], 64, nullptr, 0}, nimg, H2, W2);
    conv_forward_lin(N, 10, Src{a[3], 32, a[9], 64}, nimg, H1, W1);
    conv_forward_lin(N, 11, Src{a[10], 32, nullptr, 0}, nimg, H1, W1);
    conv_forward_lin(N, 12, Src{a[1], 16, a[11], 32}, nimg, H, W);
    conv_forward_lin(N, 13, Src{a[12], 16, nullptr, 0}, nimg, H, W);
    const int64_t P = static_cast<int64_t>(nimg) * H * W;
    hipLaunchKernelGGL(k_out_fwd, dim3(blocks_for(P, 256, 8192)), dim3(256), 0, s, a[13], N->w32 + N->L.out_w,
                       static_cast<const float*>(nullptr), N->y, P);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

// accumulates parameter gradients of the loss with dL/dy = N->dy, after unet_forward
// (linear = false) or after unet_forward + unet_forward_linear (linear = true: inputs from the
// linear pass, ReLU derivatives from the ReLU pass, no bias gradients)
int unet_backward(asd2_unet* N, int nimg, bool linear) {
    const int H = N->H, W = N->W;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
    const int64_t P0 = static_cast<int64_t>(nimg) * H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    hipStream_t s = N->stream;
    uint16_t** a = linear ? N->lin : N->act;   // inputs of each conv
    uint16_t** m = N->act;                     // ReLU patterns
    uint16_t** pooled = linear ? N->linpooled : N->pooled;
    const uint16_t* fin = linear ? N->flin : N->feat;
    const bool bias = !linear;
    // scratch layout inside g0 (level 0), g1 (levels 1-3), g2 (concat inputs)
    uint16_t* d13 = N->g0;                 // [P0][16]
    uint16_t* d12 = N->g0 + P0 * 16;       // [P0][16]
    uint16_t* dX48 = N->g2;                // [P0][48]
    uint16_t* d11 = N->g1;                 // [P1][32]
    uint16_t* d10 = d11 + P1 * 32;         // [P1][32]
    uint16_t* dX96 = N->g2 + P0 * 48;      // [P1][96]
    uint16_t* d9 = d10 + P1 * 32;          // [P2][64]
    uint16_t* d8 = d9 + P2 * 64;           // [P2][64]
    uint16_t* dX192 = dX96 + P1 * 96;      // [P2][192]
    uint16_t* d7 = d8 + P2 * 64;           // [P3][128]
    uint16_t* d6 = d7 + P3 * 128;          // [P3][128]
    uint16_t* dp2 = d6 + P3 * 128;         // [P3][64]
    uint16_t* d5 = dp2 + P3 * 64;          // [P2][64]
    uint16_t* d4 = d5 + P2 * 64;           // [P2][64]
    uint16_t* dp1 = d4 + P2 * 64;          // [P2][32]
    uint16_t* d3 = dp1 + P2 * 32;          // [P1][32]
    uint16_t* d2 = d3 + P1 * 32;           // [P1][32]
    uint16_t* dp0 = d2 + P1 * 32;          // [P1][16]
    uint16_t* d1 = N->g0 + P0 * 32;        // [P0][16]
    uint16_t* d0 = N->g0 + P0 * 48;        // [P0][16]

    {  // output layer: 16 weights and the bias (stored after them) from block partials
        const int nb = blocks_for(P0, 256, 2048);
        hipLaunchKernelGGL(k_out_bwd, dim3(nb), dim3(256), 0, s, a[13], m[13], N->w32 + N->L.out_w, N->dy, d13,
                           N->gpart, static_cast<float*>(nullptr), P0);
        sum_blocks(s, N->gpart, nb, bias ? 17 : 16, N->grad + N->L.out_w, 17);
    }
    // dec0
    conv_wgrad(N, 13, Src{a[12], 16, nullptr, 0}, d13, nimg, H, W, bias);
    conv_dgrad(N, 13, d13, d12, m[12], nimg, H, W);
    conv_wgrad(N, 12, Src{a[1], 16, a[11], 32}, d12, nimg, H, W, bias);
    conv_dgrad(N, 12, d12, dX48, nullptr, nimg, H, W);
    hipLaunchKernelGGL(k_unconcat_deep, dim3(blocks_for(P1 * 8)), dim3(256), 0, s, dX48, 16, 32, m[11], d11, nimg, H, W);
    // dec1
    conv_wgrad(N, 11, Src{a[10], 32, nullptr, 0}, d11, nimg, H1, W1, bias);
    conv_dgrad(N, 11, d11, d10, m[10], nimg, H1, W1);
    conv_wgrad(N, 10, Src{a[3], 32, a[9], 64}, d10, nimg, H1, W1, bias);
    conv_dgrad(N, 10, d10, dX96, nullptr, nimg, H1, W1);
    hipLaunchKernelGGL(k_unconcat_deep, dim3(blocks_for(P2 * 16)), dim3(256), 0, s, dX96, 32, 64, m[9], d9, nimg, H1,
                       W1);
    // dec2
    conv_wgrad(N, 9, Src{a[8], 64, nullptr, 0}, d9, nimg, H2, W2, bias);
    conv_dgrad(N, 9, d9, d8, m[8], nimg, H2, W2);
    conv_wgrad(N, 8, Src{a[5], 64, a[7], 128}, d8, nimg, H2, W2, bias);
    conv_dgrad(N, 8, d8, dX192, nullptr, nimg, H2, W2);
    hipLaunchKernelGGL(k_unconcat_deep, dim3(blocks_for(P3 * 32)), dim3(256), 0, s, dX192, 64, 128, m[7], d7, nimg, H2,
                       W2);
    // enc3
    conv_wgrad(N, 7, Src{a[6], 128, nullptr, 0}, d7, nimg, H3, W3, bias);
    conv_dgrad(N, 7, d7, d6, m[6], nimg, H3, W3);
    conv_wgrad(N, 6, Src{pooled[2], 64, nullptr, 0}, d6, nimg, H3, W3, bias);
    conv_dgrad(N, 6, d6, dp2, nullptr, nimg, H3, W3);
    // enc2
    hipLaunchKernelGGL(k_skip_pool_mask, dim3(blocks_for(P2 * 16)), dim3(256), 0, s, dX192, 192, dp2, m[5], d5, nimg,
                       H2, W2, 64);
    conv_wgrad(N, 5, Src{a[4], 64, nullptr, 0}, d5, nimg, H2, W2, bias);
    conv_dgrad(N, 5, d5, d4, m[4], nimg, H2, W2);
    conv_wgrad(N, 4, Src{pooled[1], 32, nullptr, 0}, d4, nimg, H2, W2, bias);
    conv_dgrad(N, 4, d4, dp1, nullptr, nimg, H2, W2);
    // enc1
    hipLaunchKernelGGL(k_skip_pool_mask, dim3(blocks_for(P1 * 8)), dim3(256), 0, s, dX96, 96, dp1, m[3], d3, nimg, H1,
                       W1, 32);
    conv_wgrad(N, 3, Src{a[2], 32, nullptr, 0}, d3, nimg, H1, W1, bias);
    conv_dgrad(N, 3, d3, d2, m[2], nimg, H1, W1);
    conv_wgrad(N, 2, Src{pooled[0], 16, nullptr, 0}, d2, nimg, H1, W1, bias);
    conv_dgrad(N, 2, d2, dp0, nullptr, nimg, H1, W1);
    // enc0
    hipLaunchKernelGGL(k_skip_pool_mask, dim3(blocks_for(P0 * 4)), dim3(256), 0, s, dX48, 48, dp0, m[1], d1, nimg, H, W,
                       16);
    conv_wgrad(N, 1, Src{a[0], 16, nullptr, 0}, d1, nimg, H, W, bias);
    conv_dgrad(N, 1, d1, d0, m[0], nimg, H, W);
    conv_wgrad(N, 0, Src{fin, 4, nullptr, 0}, d0, nimg, H, W, bias);
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

}  // namespace asd2

using namespace asd2;

extern "C" {

int asd2_unet_create(int nx, int ny, int coarsen, int max_batch, unsigned long long seed, asd2_unet** out) {
    if (!out) { set_error("null argument"); return ASD2_ERR_ARG; }
    *out = nullptr;
    const bool pow2 = coarsen >= 1 && coarsen <= 32 && (coarsen & (coarsen - 1)) == 0;
    if (!pow2 || nx <= 0 || ny <= 0 || nx % (8 * coarsen) || ny % (8 * coarsen) || max_batch <= 0) {
        set_error("the network needs a power-of-two coarsening c <= 32, nx and ny divisible by 8 c and a positive "
                  "batch (got %d x %d, c = %d, batch %d)", nx, ny, coarsen, max_batch);
        return ASD2_ERR_UNSUPPORTED;
    }
    auto* N = new asd2_unet;
    N->H = nx / coarsen;
    N->W = ny / coarsen;
    N->coarsen = coarsen;
    N->max_batch = max_batch;
    N->L = unet_layout();
    auto fail = [&](int st) { asd2_unet_destroy(N); return st; };
    if (hipStreamCreateWithFlags(&N->own_stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(ASD2_ERR_HIP);
    }
    N->stream = N->own_stream;
    const int64_t T = N->L.total;
    const int64_t P0 = static_cast<int64_t>(max_batch) * N->H * N->W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    const int64_t CS = static_cast<int64_t>(max_batch) * (N->H + 1) * (N->W + 1);
    int st = ASD2_OK;
    if ((st = dalloc(&N->w32, T)) || (st = dalloc(&N->grad, T)) || (st = dalloc(&N->adam_m, T)) ||
        (st = dalloc(&N->adam_v, T)) || (st = dalloc(&N->wf, T)) || (st = dalloc(&N->wb, T)) ||
        (st = dalloc(&N->feat, P0 * 4)) || (st = dalloc(&N->y, P0)) || (st = dalloc(&N->dy, P0)) ||
        (st = dalloc(&N->scratch, 8LL * max_batch + 64)) || (st = dalloc(&N->partials, 4096)) ||
        (st = dalloc(&N->gpart, N->gpart_n = gpart_size(max_batch))) ||
        (st = dalloc(&N->tg, P0)) || (st = dalloc(&N->tacx, CS)) || (st = dalloc(&N->tacy, CS)) ||
        (st = dalloc(&N->tdc, CS)))
        return fail(st);
    const int64_t lvl[4] = {P0, P1, P2, P3};
    for (int l = 0; l < UNET_NCONV; l++)
        if ((st = dalloc(&N->act[l], lvl[UNET_CONVS[l].level] * UNET_CONVS[l].cout)) ||
            (st = dalloc(&N->lin[l], lvl[UNET_CONVS[l].level] * UNET_CONVS[l].cout)))
            return fail(st);
    if ((st = dalloc(&N->pooled[0], P1 * 16)) || (st = dalloc(&N->pooled[1], P2 * 32)) ||
        (st = dalloc(&N->pooled[2], P3 * 64)) || (st = dalloc(&N->linpooled[0], P1 * 16)) ||
        (st = dalloc(&N->linpooled[1], P2 * 32)) || (st = dalloc(&N->linpooled[2], P3 * 64)) ||
        (st = dalloc(&N->flin, P0 * 4)))
        return fail(st);
    if ((st = dalloc(&N->g0, P0 * 64)) || (st = dalloc(&N->g1, P1 * 32 * 4 + P2 * 64 * 5 + P3 * (128 * 2 + 64) + P1 * 16)) ||
        (st = dalloc(&N->g2, P0 * 48 + P1 * 96 + P2 * 192)))
        return fail(st);
    // He-uniform initial weights from a splitmix64 stream (host), zero biases
    std::vector<float> h(static_cast<size_t>(T), 0.0f);
    uint64_t x = seed + 0x9E3779B97F4A7C15ull;
    auto next = [&]() {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < UNET_NCONV; l++) {
        const ConvSpec& c = UNET_CONVS[l];
        const int cin = c.cin_a + c.cin_b;
        const double lim = std::sqrt(6.0 / (9.0 * cin));
        for (int64_t e = 0; e < 9LL * cin * c.cout; e++) h[N->L.w[l] + e] = static_cast<float>((2 * next() - 1) * lim);
    }
    // the output layer too: the loss is invariant to the scale of y, and y = 0 has no gradient
    const int c0 = UNET_CONVS[UNET_NCONV - 1].cout;
    for (int c = 0; c < c0; c++) h[N->L.out_w + c] = static_cast<float>((2 * next() - 1) * std::sqrt(6.0 / c0));
    if (hipMemcpy(N->w32, h.data(), T * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("upload of the initial weights failed");
        return fail(ASD2_ERR_HIP);
    }
    if ((st = unet_refresh(N))) return fail(st);
    if (hipStreamSynchronize(N->stream) != hipSuccess) {
        set_error("network set-up failed");
        return fail(ASD2_ERR_HIP);
    }
    *out = N;
    return ASD2_OK;
}

int asd2_unet_destroy(asd2_unet* N) {
    if (!N) return ASD2_OK;
    if (N->own_stream) (void)hipStreamSynchronize(N->own_stream);
    void* bufs[] = {N->w32, N->grad, N->adam_m, N->adam_v, N->wf, N->wb, N->feat, N->y, N->dy, N->scratch,
                    N->partials, N->pooled[0], N->pooled[1], N->pooled[2], N->g0, N->g1, N->g2, N->tg, N->tacx,
                    N->tacy, N->tdc, N->flin, N->linpooled[0], N->linpooled[1], N->linpooled[2], N->gpart};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    for (auto* p : N->act)
        if (p) (void)hipFree(p);
    for (auto* p : N->lin)
        if (p) (void)hipFree(p);
    if (N->own_stream) (void)hipStreamDestroy(N->own_stream);
    delete N;
    return ASD2_OK;
}

int asd2_unet_param_count(const asd2_unet* N, long long* count) {
    ASD2_TRY(check_net(N));
    if (!count) { set_error("null argument"); return ASD2_ERR_ARG; }
    *count = N->L.total;
    return ASD2_OK;
}

int asd2_unet_set_params(asd2_unet* N, const float* host, long long count) {
    ASD2_TRY(check_net(N));
    if (!host || count != N->L.total) { set_error("expected %lld parameters", static_cast<long long>(N->L.total)); return ASD2_ERR_ARG; }
    ASD2_HIP(hipMemcpyAsync(N->w32, host, count * sizeof(float), hipMemcpyHostToDevice, N->stream));
    ASD2_TRY(unet_refresh(N));
    ASD2_HIP(hipStreamSynchronize(N->stream));
    return ASD2_OK;
}

int asd2_unet_get_params(asd2_unet* N, float* host, long long count) {
    ASD2_TRY(check_net(N));
    if (!host || count != N->L.total) { set_error("expected %lld parameters", static_cast<long long>(N->L.total)); return ASD2_ERR_ARG; }
    ASD2_HIP(hipMemcpyAsync(host, N->w32, count * sizeof(float), hipMemcpyDeviceToHost, N->stream));
    ASD2_HIP(hipStreamSynchronize(N->stream));
    return ASD2_OK;
}

int asd2_unet_get_grads(asd2_unet* N, float* host, long long count) {
    ASD2_TRY(check_net(N));
    if (!host || count != N->L.total) { set_error("expected %lld parameters", static_cast<long long>(N->L.total)); return ASD2_ERR_ARG; }
    ASD2_HIP(hipMemcpyAsync(host, N->grad, count * sizeof(float), hipMemcpyDeviceToHost, N->stream));
    ASD2_HIP(hipStreamSynchronize(N->stream));
    return ASD2_OK;
}

// features f (FP32 [B][nx][ny][4], device) -> y (FP32 [B][nx][ny], device); keeps activations
int asd2_unet_forward(asd2_unet* N, int batch, const float* d_feat, float* d_y) {
    ASD2_TRY(check_net(N));
    if (batch <= 0 || batch > N->max_batch || !d_feat || !d_y) { set_error("bad argument"); return ASD2_ERR_ARG; }
    const int64_t P = static_cast<int64_t>(batch) * N->H * N->W;
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks_for(P * 4, 256, 8192)), dim3(256), 0, N->stream, d_feat, N->feat,
                       P * 4);
    ASD2_TRY(unet_forward(N, batch));
    ASD2_HIP(hipMemcpyAsync(d_y, N->y, P * sizeof(float), hipMemcpyDeviceToDevice, N->stream));
    ASD2_HIP(hipStreamSynchronize(N->stream));
    return ASD2_OK;
}

// linear pass: masks from the ReLU pass on the coefficient channels (channel 0 zeroed), then the
// masked pass on channel 0 alone; y = FP32 [B][H][W]
int asd2_unet_forward_linear(asd2_unet* N, int batch, const float* d_feat, float* d_y) {
    ASD2_TRY(check_net(N));
    if (batch <= 0 || batch > N->max_batch || !d_feat || !d_y) { set_error("bad argument"); return ASD2_ERR_ARG; }
    const int64_t P = static_cast<int64_t>(batch) * N->H * N->W;
    hipLaunchKernelGGL(k_split_features, dim3(blocks_for(P, 256, 8192)), dim3(256), 0, N->stream, d_feat, N->feat,
                       N->flin, P);
    ASD2_TRY(unet_forward(N, batch));
    ASD2_TRY(unet_forward_linear(N, batch));
    ASD2_HIP(hipMemcpyAsync(d_y, N->y, P * sizeof(float), hipMemcpyDeviceToDevice, N->stream));
    ASD2_HIP(hipStreamSynchronize(N->stream));
    return ASD2_OK;
}

// parameter gradients of the linear pass for dL/dy = d_dy (after asd2_unet_forward_linear)
int asd2_unet_backward_linear(asd2_unet* N, int batch, const float* d_dy) {
    ASD2_TRY(check_net(N));
    if (batch <= 0 || batch > N->max_batch || !d_dy) { set_error("bad argument"); return ASD2_ERR_ARG; }
    const int64_t P = static_cast<int64_t>(batch) * N->H * N->W;
    ASD2_HIP(hipMemsetAsync(N->grad, 0, N->L.total * sizeof(float), N->stream));
    ASD2_HIP(hipMemcpyAsync(N->dy, d_dy, P * sizeof(float), hipMemcpyDeviceToDevice, N->stream));
    ASD2_TRY(unet_backward(N, batch, true));
    ASD2_HIP(hipStreamSynchronize(N->stream));
    return ASD2_OK;
}

// parameter gradients for dL/dy = d_dy (after asd2_unet_forward on the same batch)
int asd2_unet_backward(asd2_unet* N, int batch, const float* d_dy) {
    ASD2_TRY(check_net(N));
    if (batch <= 0 || batch > N->max_batch || !d_dy) { set_error("bad argument"); return ASD2_ERR_ARG; }
    const int64_t P = static_cast<int64_t>(batch) * N->H * N->W;
    ASD2_HIP(hipMemsetAsync(N->grad, 0, N->L.total * sizeof(float), N->stream));
    ASD2_HIP(hipMemcpyAsync(N->dy, d_dy, P * sizeof(float), hipMemcpyDeviceToDevice, N->stream));
    ASD2_TRY(unet_backward(N, batch, false));
    ASD2_HIP(hipStreamSynchronize(N->stream));
    return ASD2_OK;
}

// One Adam step on the energy loss of `batch` coarse systems (residual g, the level's operator, and
// optionally the energy g.A_c^-1 g of each sample).
int asd2_unet_train_step_coarse(asd2_unet* N, int batch, const float* d_g, const float* d_acx, const float* d_acy,
                                const float* d_dc, const float* d_energy, double h2x, double h2y, float lr,
                                float* loss_out) {
    ASD2_TRY(check_net(N));
    if (batch <= 0 || batch > N->max_batch || !d_g || !d_acx || !d_acy || !d_dc || !(lr > 0.f) || !(h2x > 0.0) ||
        !(h2y > 0.0)) {
        set_error("bad argument");
        return ASD2_ERR_ARG;
    }
    const int Hc = N->H, Wc = N->W;
    const int64_t Pc = static_cast<int64_t>(Hc) * Wc;
    hipStream_t s = N->stream;
    float* sums = N->scratch;  // [TS B]
    float* sp = N->gpart;      // block partials of the sums
    ASD2_HIP(hipMemsetAsync(sums, 0, TS * batch * sizeof(float), s));
    ASD2_HIP(hipMemsetAsync(N->grad, 0, N->L.total * sizeof(float), s));
    const dim3 gc(blocks_for(Pc, 256, 1024), batch);
    const int nbc = static_cast<int>(gc.x);
    hipLaunchKernelGGL(k_tr_gnorm, gc, dim3(256), 0, s, d_g, d_dc, Hc, Wc, sp);
    hipLaunchKernelGGL(k_tr_finish, dim3(1), dim3(64), 0, s, sp, nbc, batch, 0, 3, d_energy, sums);
    hipLaunchKernelGGL(k_tr_features, gc, dim3(256), 0, s, d_g, d_acx, d_acy, d_dc, Hc, Wc, static_cast<float>(h2x),
                       static_cast<float>(h2y), sums, N->feat, N->flin);
    ASD2_TRY(unet_forward(N, batch));
    ASD2_TRY(unet_forward_linear(N, batch));
    hipLaunchKernelGGL(k_tr_gamma, gc, dim3(256), 0, s, N->y, d_g, d_acx, d_acy, d_dc, Hc, Wc, sp);
    hipLaunchKernelGGL(k_tr_finish, dim3(1), dim3(64), 0, s, sp, nbc, batch, 3, 2, static_cast<const float*>(nullptr),
                       sums);
    hipLaunchKernelGGL(k_tr_dy, gc, dim3(256), 0, s, N->y, d_g, d_acx, d_acy, d_dc, Hc, Wc, sums, N->dy, batch);
    ASD2_TRY(unet_backward(N, batch, true));
    N->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, static_cast<float>(N->adam_t)), c2 = 1.f - std::pow(b2, static_cast<float>(N->adam_t));
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(N->L.total, 256, 4096)), dim3(256), 0, s, N->w32, N->grad, N->adam_m,
                       N->adam_v, N->L.total, lr, b1, b2, 1e-8f, c1, c2, 1.0f);
    ASD2_TRY(unet_refresh(N));
    std::vector<float> h(static_cast<size_t>(TS * batch));
    ASD2_HIP(hipMemcpyAsync(h.data(), sums, h.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    ASD2_HIP(hipStreamSynchronize(s));
    if (loss_out) {
        // per sample -(y.g)^2 / (y.A_c y) / nrm: the energy of the best step along the network's
        // correction relative to the sample's energy (1 + that: the A_c-norm error left, squared and
        // relative) or to its Jacobi scale
        double L = d_energy ? static_cast<double>(batch) : 0.0;
        for (int b = 0; b < batch; b++) {
            const double nrm = h[TS * b], yg = h[TS * b + 3], yay = h[TS * b + 4];
            if (nrm > 0.0 && yay > 0.0) L -= yg * yg / yay / nrm;
        }
        *loss_out = static_cast<float>(L / batch);
    }
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

int asd2_unet_train_step(asd2_unet* N, int batch, const float* d_r, const float* d_dinv, const float* d_ax,
                         const float* d_ay, const float* d_dg, double dx, double dy, float lr, float* loss_out) {
    ASD2_TRY(check_net(N));
    if (batch <= 0 || batch > N->max_batch || !d_r || !d_dinv || !d_ax || !d_ay || !d_dg || !(lr > 0.f)) {
        set_error("bad argument");
        return ASD2_ERR_ARG;
    }
    const int c = N->coarsen, Hc = N->H, Wc = N->W, nx = Hc * c, ny = Wc * c;
    const int64_t stride = static_cast<int64_t>(nx + 6) * (ny + 6);
    const int64_t Pf = static_cast<int64_t>(nx) * ny, Pc = static_cast<int64_t>(Hc) * Wc;
    const int64_t cs = static_cast<int64_t>(Hc + 1) * (Wc + 1);
    hipStream_t s = N->stream;
    float* sums = N->scratch;  // [TS B]
    float* sp = N->gpart;      // block partials of the sums
    ASD2_HIP(hipMemsetAsync(sums, 0, TS * batch * sizeof(float), s));
    ASD2_HIP(hipMemsetAsync(N->grad, 0, N->L.total * sizeof(float), s));
    hipLaunchKernelGGL(k_tr_coarse_op, dim3(blocks_for(cs, 256, 1024), batch), dim3(256), 0, s, d_ax, d_ay, d_dg,
                       stride, ny + 6, c, Hc, Wc, N->tacx, N->tacy, N->tdc);
    const dim3 gc(blocks_for(Pc, 256, 1024), batch);
    const int nbc = static_cast<int>(gc.x);
    hipLaunchKernelGGL(k_tr_restrict, gc, dim3(256), 0, s, d_r, d_dinv, d_ax, d_ay, d_dg, stride, nx, ny, c, N->tg, sp);
    hipLaunchKernelGGL(k_tr_finish, dim3(1), dim3(64), 0, s, sp, nbc, batch, 0, 2, static_cast<const float*>(nullptr),
                       sums);
    hipLaunchKernelGGL(k_tr_usum, gc, dim3(256), 0, s, N->tg, N->tdc, Hc, Wc, sp);
    hipLaunchKernelGGL(k_tr_finish, dim3(1), dim3(64), 0, s, sp, nbc, batch, 2, 1, static_cast<const float*>(nullptr),
                       sums);
    hipLaunchKernelGGL(k_tr_features, gc, dim3(256), 0, s, N->tg, N->tacx, N->tacy, N->tdc, Hc, Wc,
                       static_cast<float>(dx * dx / c), static_cast<float>(dy * dy / c), sums, N->feat, N->flin);
    ASD2_TRY(unet_forward(N, batch));
    ASD2_TRY(unet_forward_linear(N, batch));
    hipLaunchKernelGGL(k_tr_gamma, gc, dim3(256), 0, s, N->y, N->tg, N->tacx, N->tacy, N->tdc, Hc, Wc, sp);
    hipLaunchKernelGGL(k_tr_finish, dim3(1), dim3(64), 0, s, sp, nbc, batch, 3, 2, static_cast<const float*>(nullptr),
                       sums);
    hipLaunchKernelGGL(k_tr_dy, gc, dim3(256), 0, s, N->y, N->tg, N->tacx, N->tacy, N->tdc, Hc, Wc, sums, N->dy,
                       batch);
    ASD2_TRY(unet_backward(N, batch, true));
    N->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, static_cast<float>(N->adam_t)), c2 = 1.f - std::pow(b2, static_cast<float>(N->adam_t));
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(N->L.total, 256, 4096)), dim3(256), 0, s, N->w32, N->grad, N->adam_m,
                       N->adam_v, N->L.total, lr, b1, b2, 1e-8f, c1, c2, 1.0f);
    ASD2_TRY(unet_refresh(N));
    std::vector<float> h(static_cast<size_t>(TS * batch));
    ASD2_HIP(hipMemcpyAsync(h.data(), sums, h.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    ASD2_HIP(hipStreamSynchronize(s));
    if (loss_out) {
        // per sample (E(D^-1 r) - (y.g)^2 / y.A_c y) / r.D^-1 r with E(D^-1 r) = -(r.D^-1 r) - (D^-1 r).w
        double L = 0.0;
        for (int b = 0; b < batch; b++) {
            const double cn = h[TS * b], ej = h[TS * b + 1], yg = h[TS * b + 3], yay = h[TS * b + 4];
            if (!(cn > 0.0)) continue;
            double e = -cn - ej;
            if (yay > 0.0) e -= yg * yg / yay;
            L += e / cn;
        }
        *loss_out = static_cast<float>(L / batch);
    }
    ASD2_LAUNCH_CHECK();
    return ASD2_OK;
}

}  // extern "C"
