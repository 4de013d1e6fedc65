// Stable LSD radix sort of (uint32 key, int32 value) pairs for the LBVH key
// sort (replaces thrust::sort_by_key, bvh.cu:795-807 / 888-894).
//
// One pass per 8-bit digit, three kernels per pass, over tiles of
// kSortTile keys:
//   hist    — each 512-thread block counts the digits of its tile; lanes
//             with equal digits are found with 8 wave ballots, so only one
//             lane per (wave, digit) touches LDS.
//   scan    — one block per digit turns its row of tile counts into an
//             exclusive prefix and records the digit total.
//   scatter — each block sorts its tile by the digit in LDS, then writes it
//             out: the tile's keys of one digit go to one contiguous run of
//             the output, so consecutive threads write consecutive
//             addresses.  A key's place in its digit's run: keys of that
//             digit in earlier rounds + earlier waves + earlier lanes of
//             its wave (ballot match), so equal digits keep their input
//             order and the sort is stable.
#include "arc_internal.hpp"

namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr int kTile = arc::kSortTile;
constexpr int kItems = kTile / kThreads;
constexpr int kRadix = arc::kSortRadix;
constexpr int kScanThreads = 256;
static_assert(kTile % kThreads == 0 && kRadix <= kThreads, "sort tile shape");

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
    if (tid < kRadix) h[tid] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const int idx = base + j * kThreads + tid;
        const bool valid = idx < n;
        const uint32_t digit = valid ? (keys[idx] >> shift) & 0xffu : 0u;
        const uint64_t m = match_digit(__ballot(valid), digit);
        const uint64_t lower = m & ((1ull << lane) - 1ull);
        if (valid && lower == 0ull) atomicAdd(&h[digit], (uint32_t)__popcll(m));
    }
    __syncthreads();
    if (tid < kRadix) counts[(size_t)tid * num_tiles + blockIdx.x] = h[tid];
}

// Exclusive scan of sh[0, kRadix) in place, by the whole block
// (Hillis-Steele over the first kRadix threads; every thread reaches every
// barrier).  Returns the total.
__device__ inline uint32_t radix_exclusive_scan(uint32_t* sh, uint32_t* tmp) {
    const int tid = threadIdx.x;
    uint32_t v = 0;
    if (tid < kRadix) {
        v = sh[tid];
        tmp[tid] = v;
    }
    __syncthreads();
    for (int off = 1; off < kRadix; off <<= 1) {
        uint32_t u = 0;
        if (tid < kRadix && tid >= off) u = tmp[tid - off];
        __syncthreads();
        if (tid < kRadix) tmp[tid] += u;
        __syncthreads();
    }
    const uint32_t total = tmp[kRadix - 1];
    if (tid < kRadix) sh[tid] = tmp[tid] - v;
    __syncthreads();
    return total;
}

// Block d scans row d of the digit-major count table (exclusive, in place)
// and writes the digit's total count.
__global__ __launch_bounds__(kScanThreads) void k_scan_digit(uint32_t* __restrict__ counts, int num_tiles,
                                                            uint32_t* __restrict__ totals) {
    __shared__ uint32_t sh[kScanThreads];
    const int tid = threadIdx.x;
    uint32_t* row = counts + (size_t)blockIdx.x * num_tiles;
    uint32_t run = 0;
    for (int base = 0; base < num_tiles; base += kScanThreads) {
        const int t = base + tid;
        const uint32_t v = t < num_tiles ? row[t] : 0u;
        sh[tid] = v;
        __syncthreads();
        for (int off = 1; off < kScanThreads; off <<= 1) {
            const uint32_t u = tid >= off ? sh[tid - off] : 0u;
            __syncthreads();
            sh[tid] += u;
            __syncthreads();
        }
        if (t < num_tiles) row[t] = run + sh[tid] - v;
        run += sh[kScanThreads - 1];
        __syncthreads();
    }
    if (tid == 0) totals[blockIdx.x] = run;
}

__global__ __launch_bounds__(kThreads) void k_scatter(const uint32_t* __restrict__ keys_in,
                                                      const int* __restrict__ vals_in,
                                                      uint32_t* __restrict__ keys_out,
                                                      int* __restrict__ vals_out, int n, int shift,
                                                      const uint32_t* __restrict__ offsets,
                                                      const uint32_t* __restrict__ totals, int num_tiles) {
    __shared__ uint32_t s_keys[kTile];
    __shared__ int s_vals[kTile];
    __shared__ uint32_t local_start[kRadix];  // the tile's digit counts, then their exclusive scan
    __shared__ uint32_t out_start[kRadix];    // where the tile's run of each digit starts in the output
    __shared__ uint32_t run[kRadix];          // keys of each digit placed in earlier rounds
    __shared__ uint32_t wcount[kWaves][kRadix];
    __shared__ uint32_t tmp[kRadix];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const int tile = blockIdx.x;
    const int base = tile * kTile;
    const int tile_n = min(kTile, n - base);
    // digit bases: the exclusive scan of the digit totals, plus this tile's
    // offset within each digit (the scanned count table); the tile's own
    // count of digit d is the next tile's offset minus its own
    if (tid < kRadix) {
        const uint32_t off = offsets[(size_t)tid * num_tiles + tile];
        const uint32_t next = tile + 1 < num_tiles ? offsets[(size_t)tid * num_tiles + tile + 1] : totals[tid];
        local_start[tid] = next - off;
        out_start[tid] = totals[tid];
        run[tid] = off;
    }
    __syncthreads();
    radix_exclusive_scan(out_start, tmp);  // out_start = digit bases
    radix_exclusive_scan(local_start, tmp);
    if (tid < kRadix) {
        out_start[tid] += run[tid];
        run[tid] = 0;
    }
    // 1. the tile sorted by digit into LDS, stably, a round of kThreads keys at a time
    for (int j = 0; j < kItems; ++j) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w)
            if (tid < kRadix) wcount[w][tid] = 0;
        __syncthreads();
        const int i = j * kThreads + tid;
        const bool valid = i < tile_n;
        const uint32_t key = valid ? keys_in[base + i] : 0u;
        const int val = valid ? vals_in[base + i] : 0;
        const uint32_t digit = (key >> shift) & 0xffu;
        const uint64_t m = match_digit(__ballot(valid), digit);
        const uint64_t lower = m & ((1ull << lane) - 1ull);
        if (valid && lower == 0ull) wcount[wave][digit] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pos = local_start[digit] + run[digit] + (uint32_t)__popcll(lower);
            for (int w = 0; w < wave; ++w) pos += wcount[w][digit];
            s_keys[pos] = key;
            s_vals[pos] = val;
        }
        __syncthreads();
        if (tid < kRadix) {
            uint32_t add = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) add += wcount[w][tid];
            run[tid] += add;
        }
    }
    __syncthreads();
    // 2. out in LDS order: consecutive threads, consecutive places of one run
    for (int i = tid; i < tile_n; i += kThreads) {
        const uint32_t key = s_keys[i];
        const uint32_t digit = (key >> shift) & 0xffu;
        const uint32_t pos = out_start[digit] + (uint32_t)i - local_start[digit];
        keys_out[pos] = key;
        vals_out[pos] = s_vals[i];
    }
}

}  // namespace

namespace arc {

int sort_pairs(arc_ctx* ctx, SortScratch& s, uint32_t* keys, int* vals, int n, int key_bits) {
    if (n <= 1) return ARC_OK;
    const int num_tiles = (n + kTile - 1) / kTile;
    ARC_TRY(ensure(ctx, s.keys_alt, sizeof(uint32_t) * (size_t)n));
    ARC_TRY(ensure(ctx, s.vals_alt, sizeof(int) * (size_t)n));
    ARC_TRY(ensure(ctx, s.counts, sizeof(uint32_t) * sort_counts_size((size_t)n)));
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
        hipLaunchKernelGGL(k_scan_digit, dim3(kRadix), dim3(kScanThreads), 0, ctx->stream, counts, num_tiles,
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
