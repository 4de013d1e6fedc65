// The learned cut's network on gfx950: a 24 -> 64 -> 64 -> 1 MLP (ReLU,
// sigmoid output) with BF16 weights and FP32 accumulation on the matrix cores.
//
// Layout.  Activations are kept transposed (features x samples), 16 samples
// per tile, so that each layer is D = W^T * H^T with the weights as the A
// operand and the activations as the B operand of
// v_mfma_f32_16x16x16_bf16.  For that instruction (wave64) lane l holds
//   A[l % 16][4 (l / 16) + j],  B[4 (l / 16) + j][l % 16],  D[4 (l / 16) + j][l % 16]
// for j = 0..3, so a layer's D fragment is, element for element, the next
// layer's B fragment: the hidden layers never leave the registers.  The
// weights are pre-arranged as per-lane A fragments (k_net_pack) and held in
// VGPRs for the life of a wave.
//
// Parameter vector (FP32, ARC_CUTNET_PARAMS floats, the host format):
//   [0, 24) feature mean, [24, 48) feature 1/std, W1[32][64] (rows >= 24 zero),
//   b1[64], W2[64][64] (W2[in][out]), b2[64], w3[64], b3.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "arc_internal.hpp"

namespace arc {
namespace net {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 24, kInPad = 32, kHid = 64;
constexpr int kMean = 0, kIstd = 24, kW1 = 48, kB1 = kW1 + kInPad * kHid, kW2 = kB1 + kHid,
              kB2 = kW2 + kHid * kHid, kW3 = kB2 + kHid, kB3 = kW3 + kHid, kUsed = kB3 + 1;
constexpr int kParams = 6400;
static_assert(kUsed <= kParams, "parameter vector");
// packed A fragments: W1^T (4 x 2 tiles), W2^T (4 x 4), W2 (4 x 4, backward)
constexpr int kF1 = 0, kF2 = 8, kF2b = 24, kFrags = 40;

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ inline unsigned short f2bf(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ inline bf4 pack4(float a, float b, float c, float d) {
    bf4 r;
    r.x = (short)f2bf(a);
    r.y = (short)f2bf(b);
    r.z = (short)f2bf(c);
    r.w = (short)f2bf(d);
    return r;
}
__device__ inline f4 mfma(bf4 a, bf4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// The inference weights of one lane.
struct Frags {
    bf4 w1[4][2];
    bf4 w2[4][4];
};

__device__ inline void load_frags(const bf4* __restrict__ frag, int lane, Frags& F) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) F.w1[mt][kt] = frag[(kF1 + mt * 2 + kt) * 64 + lane];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) F.w2[mt][kt] = frag[(kF2 + mt * 4 + kt) * 64 + lane];
    }
}

// This lane's eight inputs of its sample (features 4g..4g+3 and 16+4g..16+4g+3,
// g = lane / 16), normalised, as the two B fragments of layer 1.
__device__ inline void input_frags(const float* f /* [kIn] */, const float* __restrict__ params, int g, bf4& x0,
                                   bf4& x1) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        // select by the lane's group without indexing a register array at run time
        const float a0 = f[j], a1 = f[4 + j], a2 = f[8 + j], a3 = f[12 + j];
        const float c0 = f[16 + j], c1 = f[20 + j];
        v[j] = g == 0 ? a0 : g == 1 ? a1 : g == 2 ? a2 : a3;
        v[4 + j] = g == 0 ? c0 : g == 1 ? c1 : 0.f;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k0 = 4 * g + j, k1 = 16 + 4 * g + j;
        v[j] = (v[j] - params[kMean + k0]) * params[kIstd + k0];
        v[4 + j] = k1 < kIn ? (v[4 + j] - params[kMean + k1]) * params[kIstd + k1] : 0.f;
    }
    x0 = pack4(v[0], v[1], v[2], v[3]);
    x1 = pack4(v[4], v[5], v[6], v[7]);
}

// input_frags() with the normalisation read from `norm` (mean[24], then
// 1/std[24]; LDS or global).
__device__ inline void input_frags_n(const float* f, const float* norm, int g, bf4& x0, bf4& x1) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float a0 = f[j], a1 = f[4 + j], a2 = f[8 + j], a3 = f[12 + j];
        const float c0 = f[16 + j], c1 = f[20 + j];
        v[j] = g == 0 ? a0 : g == 1 ? a1 : g == 2 ? a2 : a3;
        v[4 + j] = g == 0 ? c0 : g == 1 ? c1 : 0.f;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k0 = 4 * g + j, k1 = 16 + 4 * g + j;
        v[j] = (v[j] - norm[k0]) * norm[kIn + k0];
        v[4 + j] = k1 < kIn ? (v[4 + j] - norm[k1]) * norm[kIn + k1] : 0.f;
    }
    x0 = pack4(v[0], v[1], v[2], v[3]);
    x1 = pack4(v[4], v[5], v[6], v[7]);
}

// Forward pass of 16 samples (one per lane % 16); every lane returns the
// logit of its sample.  sb: b1, b2, w3 (64 each) and b3, in LDS or global.
// All 64 lanes must be active.
__device__ inline float forward(const Frags& F, const float* sb, bf4 x0, bf4 x1, int g) {
    const f4 zero = {0.f, 0.f, 0.f, 0.f};
    bf4 h1[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f4 d = mfma(F.w1[mt][0], x0, zero);
        d = mfma(F.w1[mt][1], x1, d);
        const float* b = sb + 16 * mt + 4 * g;
        h1[mt] = pack4(fmaxf(d[0] + b[0], 0.f), fmaxf(d[1] + b[1], 0.f), fmaxf(d[2] + b[2], 0.f),
                       fmaxf(d[3] + b[3], 0.f));
    }
    float part = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f4 d = zero;
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) d = mfma(F.w2[mt][kt], h1[kt], d);
        const float* b = sb + kHid + 16 * mt + 4 * g;
        const float* w = sb + 2 * kHid + 16 * mt + 4 * g;
#pragma unroll
        for (int j = 0; j < 4; ++j) part += fmaxf(d[j] + b[j], 0.f) * w[j];
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    return part + sb[3 * kHid];
}

// forward() with the weight fragments read from LDS as they are used
// (sfrag: the first kF2b fragments, 64 lanes each), for kernels whose
// occupancy cannot spare the 48 VGPRs of Frags.
__device__ inline float forward_lds(const bf4* sfrag, const float* sb, bf4 x0, bf4 x1, int lane, int g) {
    const f4 zero = {0.f, 0.f, 0.f, 0.f};
    bf4 h1[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f4 d = mfma(sfrag[(kF1 + mt * 2) * 64 + lane], x0, zero);
        d = mfma(sfrag[(kF1 + mt * 2 + 1) * 64 + lane], x1, d);
        const float* b = sb + 16 * mt + 4 * g;
        h1[mt] = pack4(fmaxf(d[0] + b[0], 0.f), fmaxf(d[1] + b[1], 0.f), fmaxf(d[2] + b[2], 0.f),
                       fmaxf(d[3] + b[3], 0.f));
    }
    float part = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f4 d = zero;
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) d = mfma(sfrag[(kF2 + mt * 4 + kt) * 64 + lane], h1[kt], d);
        const float* b = sb + kHid + 16 * mt + 4 * g;
        const float* w = sb + 2 * kHid + 16 * mt + 4 * g;
#pragma unroll
        for (int j = 0; j < 4; ++j) part += fmaxf(d[j] + b[j], 0.f) * w[j];
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    return part + sb[3 * kHid];
}

}  // namespace net
}  // namespace arc

// The network and its training state (device memory).
struct arc_cutnet {
    arc_ctx* ctx = nullptr;
    arc::DevBuf params;  // kParams FP32
    arc::DevBuf train;   // grad, Adam m, Adam v (kParams each), loss accumulator (2 doubles)
    arc::DevBuf frag;    // kFrags x 64 lanes x bf4
    bool ready = false;  // weights set
    long long step = 0;  // Adam steps taken
};
