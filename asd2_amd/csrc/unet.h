// unet.h -- the learned preconditioner's network on gfx950 (internal).
//
// A 4-level U-Net over the coarse grid (c x c blocks of the interior), channels last (NHWC, images stacked along the
// pixel index), BF16 weights and activations, FP32 accumulation in MFMA:
//   enc0: conv3x3 4->16, 16->16 | enc1: pool, 16->32, 32->32 | enc2: pool, 32->64, 64->64
//   enc3: pool, 64->128, 128->128 | dec2: [enc2, up(enc3)] 192->64, 64->64
//   dec1: [enc1, up(dec2)] 96->32, 32->32 | dec0: [enc0, up(dec1)] 48->16, 16->16
//   out: 1x1 16->1 (linear).  Every 3x3 conv is followed by ReLU; pool = 2x2 mean,
//   up = nearest neighbour; convolutions see zeros outside the image.
// Parameters live in one FP32 vector (master copy, Adam state) with BF16 copies for the
// forward pass ([tap][c_out][c_in]) and for the input-gradient pass ([tap][c_in][c_out],
// taps mirrored).
#pragma once

#include "internal.h"

namespace asd2 {

constexpr int UNET_CIN = 4;
constexpr int UNET_NCONV = 14;

struct ConvSpec {
    int cin_a, cin_b, cout, level;  // input = [A (level res), up(B) (level + 1 res)]
};

// execution order of the 3x3 convolutions
constexpr ConvSpec UNET_CONVS[UNET_NCONV] = {
    {4, 0, 16, 0},   {16, 0, 16, 0},                  // enc0
    {16, 0, 32, 1},  {32, 0, 32, 1},                  // enc1
    {32, 0, 64, 2},  {64, 0, 64, 2},                  // enc2
    {64, 0, 128, 3}, {128, 0, 128, 3},                // enc3
    {64, 128, 64, 2}, {64, 0, 64, 2},                 // dec2
    {32, 64, 32, 1},  {32, 0, 32, 1},                 // dec1
    {16, 32, 16, 0},  {16, 0, 16, 0},                 // dec0
};

struct UNetParamLayout {
    int64_t w[UNET_NCONV], b[UNET_NCONV];  // offsets of conv weights [9][co][ci] and biases
    int64_t out_w, out_b, total;
};

UNetParamLayout unet_layout();

// BF16 conversion (round to nearest even) and the coefficient features shared by training
// and the solver (h2 = h^2 / c for the coarse coefficients of c x c blocks)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
// log10 of a face coefficient magnitude times h^2 (= 1/rho on the face), mapped to about
// [-1, 1] over density ratios up to 1000; -2 for a face the operator dropped
__device__ __forceinline__ float log_coeff(float c, float h2) {
    const float v = -c * h2;
    return v > 0.f ? (log10f(v) + 1.5f) / 1.5f : -2.0f;
}
__device__ __forceinline__ float log_diag(float d, float h2) { return (log10f(d * h2) - 0.5f) / 1.5f; }

// Coarse operator A_c = P^T A P of c x c blocks (P: piecewise-constant prolongation) from the
// FP32 padded fine arrays (gh ghost layers, idx = (i + gh) sy + j + gh for interior (i, j); ax / ay
// hold the W / S face coefficient of their cell).  Coarse arrays are (Hc + 1) x (Wc + 1),
// idx = I (Wc + 1) + J: acx = W-face coefficient of coarse cell (I, J) (row Hc: the east
// boundary faces), acy = S-face coefficient (column Wc: the north boundary faces), dc = diagonal
// (I < Hc, J < Wc): the block's fine diagonal plus twice every face inside the block.
template <typename TI, typename TO>
__device__ __forceinline__ void coarse_cell(const TI* __restrict__ ax, const TI* __restrict__ ay,
                                            const TI* __restrict__ dg, int sy, int gh, int c, int Hc, int Wc, int I,
                                            int J, TO* __restrict__ acx, TO* __restrict__ acy, TO* __restrict__ dc) {
    const int ci = I * c, cj = J * c, sc = Wc + 1;
    const int64_t base = static_cast<int64_t>(ci + gh) * sy + cj + gh;
    if (J < Wc) {
        TO s = 0;
        for (int t = 0; t < c; t++) s += static_cast<TO>(ax[base + t]);
        acx[I * sc + J] = s;
    }
    if (I < Hc) {
        TO s = 0;
        for (int t = 0; t < c; t++) s += static_cast<TO>(ay[base + static_cast<int64_t>(t) * sy]);
        acy[I * sc + J] = s;
    }
    if (I < Hc && J < Wc) {
        TO s = 0;
        for (int a = 0; a < c; a++)
            for (int b = 0; b < c; b++) {
                const int64_t k = base + static_cast<int64_t>(a) * sy + b;
                s += static_cast<TO>(dg[k]);
                if (a > 0) s += 2 * static_cast<TO>(ax[k]);
                if (b > 0) s += 2 * static_cast<TO>(ay[k]);
            }
        dc[I * sc + J] = s;
    }
}

// (A_c y)[I][J] for y row-major Hc x Wc
template <typename T, typename TY>
__device__ __forceinline__ T coarse_apply(const T* __restrict__ acx, const T* __restrict__ acy,
                                          const T* __restrict__ dc, const TY* __restrict__ y, int Hc, int Wc, int I,
                                          int J) {
    const int sc = Wc + 1, k = I * sc + J;
    const int64_t p = static_cast<int64_t>(I) * Wc + J;
    T v = dc[k] * static_cast<T>(y[p]);
    if (I > 0) v += acx[k] * static_cast<T>(y[p - Wc]);
    if (I + 1 < Hc) v += acx[k + sc] * static_cast<T>(y[p + Wc]);
    if (J > 0) v += acy[k] * static_cast<T>(y[p - 1]);
    if (J + 1 < Wc) v += acy[k + 1] * static_cast<T>(y[p + 1]);
    return v;
}

// ReLU pass of nimg images from net->feat (BF16 [nimg][H][W][4]) to net->y, on net->stream
// (set to the solver's stream while a solve uses the network); its activations are the masks
// of the linear pass
int unet_forward(::asd2_unet* net, int nimg);
// linear pass from net->flin (BF16 [nimg][H][W][4], channels 1..3 zero) to net->y: every conv
// output multiplied by (mask > 0) with the masks of the last unet_forward, no biases
// keep_pooled = false forms the pooled inputs inside the convolutions' loads (same values) and
// stores none; the backward pass needs them stored
int unet_forward_linear(::asd2_unet* net, int nimg, bool keep_pooled = true);

}  // namespace asd2

// The network object behind the C ABI (asd2_unet*).
struct asd2_unet {
    int H = 0, W = 0, max_batch = 0;      // coarse image size
    int coarsen = 1;                      // fine grid = (H c) x (W c)
    hipStream_t stream = nullptr;         // stream of the current user
    hipStream_t own_stream = nullptr;
    asd2::UNetParamLayout L{};
    float* w32 = nullptr;                 // FP32 parameters
    float* grad = nullptr;                // FP32 gradients (same layout)
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int adam_t = 0;
    uint16_t* wf = nullptr;               // BF16 forward weights (conv part of the layout)
    uint16_t* wb = nullptr;               // BF16 input-gradient weights
    uint16_t* feat = nullptr;             // [B][H][W][4]
    uint16_t* act[asd2::UNET_NCONV] = {}; // conv outputs of the ReLU pass (the masks)
    uint16_t* pooled[3] = {};             // pool(enc0), pool(enc1), pool(enc2)
    uint16_t* flin = nullptr;             // linear-pass input [B][H][W][4]
    uint16_t* lin[asd2::UNET_NCONV] = {}; // conv outputs of the linear pass
    uint16_t* linpooled[3] = {};
    uint16_t* g0 = nullptr;               // gradient scratch, largest level-0 tensor
    uint16_t* g1 = nullptr;
    uint16_t* g2 = nullptr;               // concat-input gradient scratch
    uint16_t* gs[4] = {};                 // accumulated skip gradients per level
    float* y = nullptr;                   // [B][H][W] network output
    float* dy = nullptr;                  // [B][H][W] loss gradient w.r.t. y
    float* scratch = nullptr;             // per-sample scalars of training
    float* tg = nullptr;                  // training: per sample coarse g = P^T (r - A D^-1 r)
    float* tacx = nullptr;                // training: per sample coarse operator
    float* tacy = nullptr;
    float* tdc = nullptr;
    double* partials = nullptr;
    // per-block partial sums (weight, bias and output-layer gradients, training sums), added in
    // block order by a second kernel so that every gradient is reproducible bit for bit
    float* gpart = nullptr;
    int64_t gpart_n = 0;
};
