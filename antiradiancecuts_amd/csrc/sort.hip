// Stable LSD radix sort of (uint32 key, int32 value) pairs for the LBVH key
// sort (replaces thrust::sort_by_key, bvh.cu:795-807 / 888-894).
//
// One pass per 8-bit digit, three kernels per pass:
//   hist    — each 256-thread block counts the digits of a 4096-key tile;
//             lanes with equal digits are found with 8 wave ballots so only
//             one lane per (wave, digit) touches LDS.
//   scan    — one block per digit turns its row of tile counts into an
//             exclusive prefix and records the digit total; each scatter
//             block adds the exclusive scan of the 256 totals itself.
//   scatter — each block re-reads its tile in index order, 256 keys a round;
//             rank = keys of the same digit in earlier rounds + earlier waves
//             + earlier lanes of this wave (ballot match), so equal digits
//             keep their input order: the sort is stable.
#include "arc_internal.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
constexpr int kRadix = 256;

__device__ inline uint64_t match_digit(uint64_t active, uint32_t digit) {
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (digit >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

__global__ __launch_bounds__(kThreads) void k_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                   uint32_t* __restrict__ counts, int num_tiles) {
    __shared__ uint32_t h[kRadix];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    h[tid] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
#pragma unroll 4
    for (int j = 0; j < kItems; ++j) {
        const int idx = base + j * kThreads + tid;
        const bool valid = idx < n;
        const uint32_t digit = valid ? (keys[idx] >> shift) & 0xffu : 0u;
        const uint64_t m = match_digit(__ballot(valid), digit);
        const uint64_t lower = m & ((1ull << lane) - 1ull);
        if (valid && lower == 0ull) atomicAdd(&h[digit], (uint32_t)__popcll(m));
    }
    __syncthreads();
    counts[(size_t)tid * num_tiles + blockIdx.x] = h[tid];
}

// Block-wide exclusive scan of one value per thread (kThreads threads).
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        const uint32_t u = tid >= off ? sh[tid - off] : 0u;
        __syncthreads();
        sh[tid] += u;
        __syncthreads();
    }
    const uint32_t incl = sh[tid];
    *total = sh[kThreads - 1];
    __syncthreads();
    return incl - v;
}

// Block d scans row d of the digit-major count table (exclusive, in place)
// and writes the digit's total count.
__global__ __launch_bounds__(kThreads) void k_scan_digit(uint32_t* __restrict__ counts, int num_tiles,
                                                        uint32_t* __restrict__ totals) {
    __shared__ uint32_t sh[kThreads];
    uint32_t* row = counts + (size_t)blockIdx.x * num_tiles;
    uint32_t run = 0;
    for (int base = 0; base < num_tiles; base += kThreads) {
        const int t = base + (int)threadIdx.x;
        const uint32_t v = t < num_tiles ? row[t] : 0u;
        uint32_t chunk_total;
        const uint32_t ex = block_exclusive_scan(v, sh, &chunk_total);
        if (t < num_tiles) row[t] = run + ex;
        run += chunk_total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = run;
}

__global__ __launch_bounds__(kThreads) void k_scatter(const uint32_t* __restrict__ keys_in,
                                                      const int* __restrict__ vals_in,
                                                      uint32_t* __restrict__ keys_out,
                                                      int* __restrict__ vals_out, int n, int shift,
                                                      const uint32_t* __restrict__ offsets,
                                                      const uint32_t* __restrict__ totals, int num_tiles) {
    __shared__ uint32_t base_off[kRadix];
    __shared__ uint32_t run[kRadix];
    __shared__ uint32_t wcount[kThreads / 64][kRadix];
    __shared__ uint32_t sh[kThreads];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    uint32_t all;
    const uint32_t digit_base = block_exclusive_scan(totals[tid], sh, &all);
    base_off[tid] = digit_base + offsets[(size_t)tid * num_tiles + blockIdx.x];
    run[tid] = 0;
    const int base = blockIdx.x * kTile;
    for (int j = 0; j < kItems; ++j) {
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) wcount[w][tid] = 0;
        __syncthreads();
        const int idx = base + j * kThreads + tid;
        const bool valid = idx < n;
        const uint32_t key = valid ? keys_in[idx] : 0u;
        const uint32_t digit = (key >> shift) & 0xffu;
        const uint64_t m = match_digit(__ballot(valid), digit);
        const uint64_t lower = m & ((1ull << lane) - 1ull);
        const uint32_t rank = (uint32_t)__popcll(lower);
        if (valid && lower == 0ull) wcount[wave][digit] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pos = base_off[digit] + run[digit] + rank;
            for (int w = 0; w < wave; ++w) pos += wcount[w][digit];
            keys_out[pos] = key;
            vals_out[pos] = vals_in[idx];
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) add += wcount[w][tid];
        run[tid] += add;
    }
}

}  // namespace

namespace arc {

int sort_pairs(arc_ctx* ctx, SortScratch& s, uint32_t* keys, int* vals, int n, int key_bits) {
    if (n <= 1) return ARC_OK;
    const int num_tiles = (n + kTile - 1) / kTile;
    ARC_TRY(ensure(ctx, s.keys_alt, sizeof(uint32_t) * (size_t)n));
    ARC_TRY(ensure(ctx, s.vals_alt, sizeof(int) * (size_t)n));
    ARC_TRY(ensure(ctx, s.counts, sizeof(uint32_t) * ((size_t)kRadix * num_tiles + kRadix)));
    const int passes = (key_bits + 7) / 8;
    uint32_t* kin = keys;
    int* vin = vals;
    uint32_t* kout = s.keys_alt.as<uint32_t>();
    int* vout = s.vals_alt.as<int>();
    uint32_t* counts = s.counts.as<uint32_t>();
    uint32_t* totals = counts + (size_t)kRadix * num_tiles;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_hist, dim3(num_tiles), dim3(kThreads), 0, ctx->stream, kin, n, shift,
                           counts, num_tiles);
        hipLaunchKernelGGL(k_scan_digit, dim3(kRadix), dim3(kThreads), 0, ctx->stream, counts, num_tiles,
                           totals);
        hipLaunchKernelGGL(k_scatter, dim3(num_tiles), dim3(kThreads), 0, ctx->stream, kin, vin,
                           kout, vout, n, shift, counts, totals, num_tiles);
        ARC_LAUNCHED(ctx);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys) {  // odd number of passes: copy the result back
        ARC_HIP(ctx, hipMemcpyAsync(keys, kin, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice,
                                    ctx->stream));
        ARC_HIP(ctx, hipMemcpyAsync(vals, vin, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice,
                                    ctx->stream));
    }
    return ARC_OK;
}

}  // namespace arc
