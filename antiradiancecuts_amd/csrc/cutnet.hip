// The learned cut's network: training and batch inference on gfx950.
//
// Inference of a feature table (k_net_forward) and training (k_net_train)
// use the transposed layer form of cutnet.hpp: per 16-sample tile a wave
// runs layer 1 (8 MFMAs) and layer 2 (16 MFMAs) with the weights as A
// fragments held in VGPRs.  Training adds, per tile:
//   dz        = w (sigmoid(z) - y) / sum(w)                 (lane-local)
//   dH2^T     = relu'(H2^T) * w3 dz                         (lane-local, D layout)
//   dH1^T     = relu'(H1^T) * (W2 dH2^T)                    (16 MFMAs: dH2^T is the B operand
//                                                           as it stands, W2 the A operand)
//   dW2^T    += dH2^T H1 ,  dW1^T += dH1^T X                 (24 MFMAs over the 16 samples: the
//                                                           operands are re-read transposed
//                                                           through 6 KB of LDS per wave)
// and the bias gradients as lane sums.  Each wave keeps its weight gradients
// in registers over its tiles and adds them to the global gradient once
// (FP32 atomics); k_adam then updates the FP32 master weights, clears the
// gradient, and k_net_pack re-packs the BF16 fragments.
#include <cmath>
#include <cstring>
#include <vector>

#include "cutnet.hpp"

using arc::DevBuf;
using namespace arc::net;

namespace {

constexpr int kThreads = 256;
constexpr int kTrainTiles = 8;  // 16-sample tiles per training wave

// FP32 master weights -> per-lane BF16 A fragments.
__global__ __launch_bounds__(kThreads) void k_net_pack(const float* __restrict__ P, bf4* __restrict__ frag) {
    const int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= kFrags * 64) return;
    const int f = t / 64, lane = t % 64, i = lane % 16, g = lane / 16;
    float v[4];
    for (int j = 0; j < 4; ++j) {
        if (f < kF2) {  // W1^T: A[16 mt + i][16 kt + 4 g + j] = W1[16 kt + 4 g + j][16 mt + i]
            const int mt = f / 2, kt = f % 2, in = 16 * kt + 4 * g + j;
            v[j] = in < kIn ? P[kW1 + in * kHid + 16 * mt + i] : 0.f;
        } else if (f < kF2b) {  // W2^T: A[16 mt + i][16 kt + 4 g + j] = W2[16 kt + 4 g + j][16 mt + i]
            const int mt = (f - kF2) / 4, kt = (f - kF2) % 4;
            v[j] = P[kW2 + (16 * kt + 4 * g + j) * kHid + 16 * mt + i];
        } else {  // W2: A[16 mt + i][16 kt + 4 g + j] = W2[16 mt + i][16 kt + 4 g + j]
            const int mt = (f - kF2b) / 4, kt = (f - kF2b) % 4;
            v[j] = P[kW2 + (16 * mt + i) * kHid + 16 * kt + 4 * g + j];
        }
    }
    frag[t] = pack4(v[0], v[1], v[2], v[3]);
}

__device__ inline void load_bias(const float* __restrict__ P, float* sb, int tid, int nthreads) {
    for (int j = tid; j < kHid; j += nthreads) {
        sb[j] = P[kB1 + j];
        sb[kHid + j] = P[kB2 + j];
        sb[2 * kHid + j] = P[kW3 + j];
    }
    if (tid == 0) sb[3 * kHid] = P[kB3];
}

__device__ inline void load_row(const float* __restrict__ X, int s, bool valid, float f[kIn]) {
#pragma unroll
    for (int j = 0; j < kIn; ++j) f[j] = valid ? X[(size_t)s * kIn + j] : 0.f;
}

// Probabilities of n feature rows; one wave per block, tiles strided over the grid.
__global__ __launch_bounds__(64) void k_net_forward(const float* __restrict__ P, const bf4* __restrict__ frag,
                                                    const float* __restrict__ X, int n, float* __restrict__ prob) {
    __shared__ float sb[3 * kHid + 1];
    const int lane = threadIdx.x, i = lane % 16, g = lane / 16;
    load_bias(P, sb, lane, 64);
    __syncthreads();
    Frags F;
    load_frags(frag, lane, F);
    const int tiles = (n + 15) / 16;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int s = tile * 16 + i;
        float f[kIn];
        load_row(X, s, s < n, f);
        bf4 x0, x1;
        input_frags(f, P, g, x0, x1);
        const float z = forward(F, sb, x0, x1, g);
        if (g == 0 && s < n) prob[s] = 1.f / (1.f + expf(-z));
    }
}

struct TrainBatch {
    const float* X;
    const float* y;
    const float* w;
    int n;              // rows in the table
    uint64_t a, c;      // row of permuted position k: (a k + c) mod n
    uint64_t start;     // permuted position of the batch's first row
    int count;          // rows in the batch
    float scale;        // 1 / (rows in the batch)
};

__global__ __launch_bounds__(64) void k_net_train(const float* __restrict__ P, const bf4* __restrict__ frag,
                                                  TrainBatch tb, float* __restrict__ grad,
                                                  double* __restrict__ loss) {
    __shared__ float sb[3 * kHid + 1];
    __shared__ unsigned short h1s[16][kHid], dh2s[16][kHid], dh1s[16][kHid], xs[16][kInPad];
    const int lane = threadIdx.x, i = lane % 16, g = lane / 16;
    load_bias(P, sb, lane, 64);
    __syncthreads();
    Frags F;
    load_frags(frag, lane, F);
    bf4 w2b[4][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) w2b[mt][kt] = frag[(kF2b + mt * 4 + kt) * 64 + lane];
    const f4 zero = {0.f, 0.f, 0.f, 0.f};
    f4 G2[4][4], G1[4][2];
    float gb1[4][4], gb2[4][4], gw3[4][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) G2[mt][nt] = zero;
        G1[mt][0] = G1[mt][1] = zero;
#pragma unroll
        for (int j = 0; j < 4; ++j) gb1[mt][j] = gb2[mt][j] = gw3[mt][j] = 0.f;
    }
    float gb3 = 0.f, lsum = 0.f;
    for (int tt = 0; tt < kTrainTiles; ++tt) {
        const int k = (blockIdx.x * kTrainTiles + tt) * 16 + i;
        const bool valid = k < tb.count;
        const int row = valid ? (int)((tb.a * ((tb.start + (uint64_t)k) % (uint64_t)tb.n) + tb.c) % (uint64_t)tb.n) : 0;
        float f[kIn];
        load_row(tb.X, row, valid, f);
        const float yv = valid ? tb.y[row] : 0.f;
        const float wv = valid ? (tb.w ? tb.w[row] : 1.f) : 0.f;
        bf4 x0, x1;
        input_frags(f, P, g, x0, x1);
        *reinterpret_cast<bf4*>(&xs[i][4 * g]) = x0;
        *reinterpret_cast<bf4*>(&xs[i][16 + 4 * g]) = x1;
        // forward, keeping H1^T (BF16) and H2^T (FP32)
        bf4 h1[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f4 d = mfma(F.w1[mt][0], x0, zero);
            d = mfma(F.w1[mt][1], x1, d);
            const float* b = sb + 16 * mt + 4 * g;
            h1[mt] = pack4(fmaxf(d[0] + b[0], 0.f), fmaxf(d[1] + b[1], 0.f), fmaxf(d[2] + b[2], 0.f),
                           fmaxf(d[3] + b[3], 0.f));
            *reinterpret_cast<bf4*>(&h1s[i][16 * mt + 4 * g]) = h1[mt];
        }
        float h2[4][4];
        float part = 0.f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f4 d = zero;
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) d = mfma(F.w2[mt][kt], h1[kt], d);
            const float* b = sb + kHid + 16 * mt + 4 * g;
            const float* w = sb + 2 * kHid + 16 * mt + 4 * g;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                h2[mt][j] = fmaxf(d[j] + b[j], 0.f);
                part += h2[mt][j] * w[j];
            }
        }
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        const float z = part + sb[3 * kHid];
        const float p = 1.f / (1.f + expf(-z));
        const float dz = (p - yv) * wv * tb.scale;
        if (g == 0 && valid)
            lsum += -wv * (yv * logf(fmaxf(p, 1e-7f)) + (1.f - yv) * logf(fmaxf(1.f - p, 1e-7f)));
        if (g == 0) gb3 += dz;
        bf4 dh2[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const float* w = sb + 2 * kHid + 16 * mt + 4 * g;
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                gw3[mt][j] += h2[mt][j] * dz;
                v[j] = h2[mt][j] > 0.f ? w[j] * dz : 0.f;
                gb2[mt][j] += v[j];
            }
            dh2[mt] = pack4(v[0], v[1], v[2], v[3]);
            *reinterpret_cast<bf4*>(&dh2s[i][16 * mt + 4 * g]) = dh2[mt];
        }
        // dH1^T = relu'(H1^T) * (W2 dH2^T)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f4 d = zero;
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) d = mfma(w2b[mt][kt], dh2[kt], d);
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                v[j] = h1[mt][j] > 0 ? d[j] : 0.f;  // BF16 bits of a ReLU output: > 0 iff positive
                gb1[mt][j] += v[j];
            }
            *reinterpret_cast<bf4*>(&dh1s[i][16 * mt + 4 * g]) = pack4(v[0], v[1], v[2], v[3]);
        }
        __syncthreads();  // one wave per block: orders the LDS writes above before the transposed reads
        // weight gradients over the tile's 16 samples (K = samples)
        bf4 a2[4], a1[4], bh[4], bx[2];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            a2[mt] = bf4{(short)dh2s[4 * g][16 * mt + i], (short)dh2s[4 * g + 1][16 * mt + i],
                         (short)dh2s[4 * g + 2][16 * mt + i], (short)dh2s[4 * g + 3][16 * mt + i]};
            a1[mt] = bf4{(short)dh1s[4 * g][16 * mt + i], (short)dh1s[4 * g + 1][16 * mt + i],
                         (short)dh1s[4 * g + 2][16 * mt + i], (short)dh1s[4 * g + 3][16 * mt + i]};
            bh[mt] = bf4{(short)h1s[4 * g][16 * mt + i], (short)h1s[4 * g + 1][16 * mt + i],
                         (short)h1s[4 * g + 2][16 * mt + i], (short)h1s[4 * g + 3][16 * mt + i]};
        }
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
            bx[nt] = bf4{(short)xs[4 * g][16 * nt + i], (short)xs[4 * g + 1][16 * nt + i],
                         (short)xs[4 * g + 2][16 * nt + i], (short)xs[4 * g + 3][16 * nt + i]};
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) G2[mt][nt] = mfma(a2[mt], bh[nt], G2[mt][nt]);
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) G1[mt][nt] = mfma(a1[mt], bx[nt], G1[mt][nt]);
        }
        __syncthreads();  // the next tile overwrites the LDS tiles
    }
    // bias gradients: sum over the 16 samples (lanes of one group), then one lane per group adds
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                gb1[mt][j] += __shfl_xor(gb1[mt][j], off);
                gb2[mt][j] += __shfl_xor(gb2[mt][j], off);
                gw3[mt][j] += __shfl_xor(gw3[mt][j], off);
            }
    for (int off = 1; off < 16; off <<= 1) gb3 += __shfl_xor(gb3, off);
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (i == 0) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int h = 16 * mt + 4 * g + j;
                atomicAdd(&grad[kB1 + h], gb1[mt][j]);
                atomicAdd(&grad[kB2 + h], gb2[mt][j]);
                atomicAdd(&grad[kW3 + h], gw3[mt][j]);
            }
    }
    if (lane == 0) {
        atomicAdd(&grad[kB3], gb3);
        atomicAdd(loss, (double)lsum);
    }
    // G2[mt][nt][j] = dW2^T[16 mt + 4 g + j][16 nt + i] = dW2[in 16 nt + i][out 16 mt + 4 g + j]
    // G1[mt][nt][j] = dW1^T[16 mt + 4 g + j][16 nt + i] = dW1[in 16 nt + i][hidden 16 mt + 4 g + j]
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int out = 16 * mt + 4 * g + j;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) atomicAdd(&grad[kW2 + (16 * nt + i) * kHid + out], G2[mt][nt][j]);
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const int in = 16 * nt + i;
                if (in < kIn) atomicAdd(&grad[kW1 + in * kHid + out], G1[mt][nt][j]);
            }
        }
}

// Adam on the trainable parameters (the feature normalisation is fixed);
// clears the gradient for the next step.
__global__ __launch_bounds__(kThreads) void k_adam(float* __restrict__ P, float* __restrict__ grad,
                                                   float* __restrict__ m, float* __restrict__ v, float lr,
                                                   float bc1, float bc2) {
    const int k = kW1 + blockIdx.x * kThreads + threadIdx.x;
    if (k >= kUsed) return;
    const float gk = grad[k];
    grad[k] = 0.f;
    const float mk = 0.9f * m[k] + 0.1f * gk;
    const float vk = 0.999f * v[k] + 0.001f * gk * gk;
    m[k] = mk;
    v[k] = vk;
    P[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + 1e-8f);
}

__global__ __launch_bounds__(kThreads) void k_fill(float* __restrict__ p, int n, float v) {
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k < n) p[k] = v;
}

// SplitMix64 and Box-Muller for the initial weights (host).
struct Init {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};

int upload_and_pack(arc_cutnet* net, const float* host) {
    arc_ctx* ctx = net->ctx;
    ARC_HIP(ctx, hipMemcpyAsync(net->params.p, host, sizeof(float) * kParams, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_net_pack, dim3((kFrags * 64 + kThreads - 1) / kThreads), dim3(kThreads), 0, ctx->stream,
                       net->params.as<float>(), net->frag.as<bf4>());
    ARC_LAUNCHED(ctx);
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // the host array is free to reuse on return
    net->ready = true;
    net->step = 0;
    hipLaunchKernelGGL(k_fill, dim3((3 * kParams + kThreads - 1) / kThreads), dim3(kThreads), 0, ctx->stream,
                       net->train.as<float>(), 3 * kParams, 0.f);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

}  // namespace

extern "C" {

int arc_cutnet_create(arc_ctx* ctx, arc_cutnet** out) {
    if (!ctx || !out) return ARC_E_INVALID;
    *out = nullptr;
    arc_cutnet* net = new arc_cutnet();
    net->ctx = ctx;
    int st = arc::ensure(ctx, net->params, sizeof(float) * kParams);
    if (st == ARC_OK) st = arc::ensure(ctx, net->train, sizeof(float) * 3 * kParams + 2 * sizeof(double));
    if (st == ARC_OK) st = arc::ensure(ctx, net->frag, sizeof(bf4) * kFrags * 64);
    if (st != ARC_OK) {
        arc_cutnet_destroy(net);
        return st;
    }
    *out = net;
    return ARC_OK;
}

void arc_cutnet_destroy(arc_cutnet* net) {
    if (!net) return;
    if (net->ctx) (void)hipStreamSynchronize(net->ctx->stream);
    arc::release(net->params);
    arc::release(net->train);
    arc::release(net->frag);
    delete net;
}

int arc_cutnet_init(arc_cutnet* net, uint64_t seed, const float* mean, const float* stdev) {
    if (!net || !mean || !stdev) return ARC_E_INVALID;
    std::vector<float> w(kParams, 0.f);
    for (int k = 0; k < kIn; ++k) {
        if (!std::isfinite(mean[k]) || !std::isfinite(stdev[k])) return ARC_E_INVALID;
        w[kMean + k] = mean[k];
        w[kIstd + k] = stdev[k] > 0.f ? 1.f / stdev[k] : 1.f;
    }
    Init r{seed};
    const double s1 = std::sqrt(2.0 / kIn), s2 = std::sqrt(2.0 / kHid), s3 = std::sqrt(1.0 / kHid);
    for (int in = 0; in < kIn; ++in)
        for (int h = 0; h < kHid; ++h) w[kW1 + in * kHid + h] = (float)(s1 * r.normal());
    for (int k = 0; k < kHid * kHid; ++k) w[kW2 + k] = (float)(s2 * r.normal());
    for (int k = 0; k < kHid; ++k) w[kW3 + k] = (float)(s3 * r.normal());
    return upload_and_pack(net, w.data());
}

int arc_cutnet_set_weights(arc_cutnet* net, const float* weights, int count) {
    if (!net || !weights || count != kParams) return ARC_E_INVALID;
    for (int k = 0; k < kParams; ++k)
        if (!std::isfinite(weights[k])) return arc::fail(net->ctx, ARC_E_INVALID, "non-finite weight");
    for (int in = kIn; in < kInPad; ++in)
        for (int h = 0; h < kHid; ++h)
            if (weights[kW1 + in * kHid + h] != 0.f)
                return arc::fail(net->ctx, ARC_E_INVALID, "W1 rows 24..31 must be zero");
    return upload_and_pack(net, weights);
}

int arc_cutnet_get_weights(arc_cutnet* net, float* weights, int count) {
    if (!net || !weights || count != kParams) return ARC_E_INVALID;
    if (!net->ready) return arc::fail(net->ctx, ARC_E_STATE, "cutnet without weights");
    arc_ctx* ctx = net->ctx;
    ARC_HIP(ctx, hipMemcpyAsync(weights, net->params.p, sizeof(float) * kParams, hipMemcpyDeviceToHost, ctx->stream));
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ARC_OK;
}

int arc_cutnet_forward(arc_cutnet* net, const float* X, int n, float* prob) {
    if (!net || n < 0 || (n > 0 && (!X || !prob))) return ARC_E_INVALID;
    if (!net->ready) return arc::fail(net->ctx, ARC_E_STATE, "cutnet without weights");
    if (n == 0) return ARC_OK;
    arc_ctx* ctx = net->ctx;
    const int tiles = (n + 15) / 16;
    const int grid = std::min(tiles, 4096);
    hipLaunchKernelGGL(k_net_forward, dim3(grid), dim3(64), 0, ctx->stream, net->params.as<float>(),
                       net->frag.as<bf4>(), X, n, prob);
    ARC_LAUNCHED(ctx);
    return ARC_OK;
}

int arc_cutnet_train(arc_cutnet* net, const float* X, const float* y, const float* w, int n, int epochs, int batch,
                     float learning_rate, uint64_t seed, float* loss) {
    if (!net || !X || !y || n < 1 || epochs < 0 || batch < 16 || !(learning_rate > 0.f)) return ARC_E_INVALID;
    arc_ctx* ctx = net->ctx;
    if (!net->ready) return arc::fail(ctx, ARC_E_STATE, "cutnet without weights (arc_cutnet_init)");
    float* grad = net->train.as<float>();
    float* m = grad + kParams;
    float* v = m + kParams;
    double* dloss = reinterpret_cast<double*>(v + kParams);
    const int per_wave = 16 * kTrainTiles;
    Init r{seed ^ 0x5DEECE66Dull};
    for (int e = 0; e < epochs; ++e) {
        // a permutation k -> (a k + c) mod n with gcd(a, n) = 1
        uint64_t a = 1 + r.next() % (uint64_t)n;
        while (gcd64(a, (uint64_t)n) != 1) a = 1 + (a % (uint64_t)n);
        const uint64_t c = r.next() % (uint64_t)n;
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(kThreads), 0, ctx->stream, reinterpret_cast<float*>(dloss), 2, 0.f);
        for (int start = 0; start < n; start += batch) {
            const int count = std::min(batch, n - start);
            // the loss is the batch mean: weights scale their rows' terms
            const TrainBatch tb{X, y, w, n, a, c, (uint64_t)start, count, 1.f / (float)count};
            hipLaunchKernelGGL(k_net_train, dim3((count + per_wave - 1) / per_wave), dim3(64), 0, ctx->stream,
                               net->params.as<float>(), net->frag.as<bf4>(), tb, grad, dloss);
            ++net->step;
            const float bc1 = 1.f - (float)std::pow(0.9, (double)net->step);
            const float bc2 = 1.f - (float)std::pow(0.999, (double)net->step);
            hipLaunchKernelGGL(k_adam, dim3((kUsed - kW1 + kThreads - 1) / kThreads), dim3(kThreads), 0, ctx->stream,
                               net->params.as<float>(), grad, m, v, learning_rate, bc1, bc2);
            hipLaunchKernelGGL(k_net_pack, dim3((kFrags * 64 + kThreads - 1) / kThreads), dim3(kThreads), 0,
                               ctx->stream, net->params.as<float>(), net->frag.as<bf4>());
            ARC_LAUNCHED(ctx);
        }
        if (loss) {
            double l = 0.0;
            ARC_HIP(ctx, hipMemcpyAsync(&l, dloss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
            loss[e] = (float)(l / (double)n);
        }
    }
    ARC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ARC_OK;
}

}  // extern "C"
