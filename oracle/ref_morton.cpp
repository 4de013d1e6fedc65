// Entry points over the reference's own morton.h, compiled where it lies under
// /root/reference (Framework/morton.h) by oracle/Makefile.  The result lands in
// oracle/_ref/ and is used only by tests/ to pin the oracle's Morton
// restatement (orc_morton3d / orc_morton6d / orc_interleave3 / orc_expand_*).
#include <hip/hip_runtime.h>
#include <cstdint>
#include <iostream>
#include "morton.h"

extern "C" {
uint32_t ref_expand_bits_by1(uint32_t x) { return expandBitsBy1(x); }
uint32_t ref_expand_bits_by2(uint32_t x) { return expandBitsBy2(x); }
uint32_t ref_interleave3(uint32_t x, uint32_t y, uint32_t z) { return interleave(x, y, z); }
uint64_t ref_interleave6(uint64_t x, uint64_t y, uint64_t z, uint64_t u, uint64_t v, uint64_t w) {
    return interleave(x, y, z, u, v, w);
}
uint32_t ref_morton3d(float x, float y, float z) { return MortonCode3D()(make_float3(x, y, z)); }
uint64_t ref_morton6d(float ax, float ay, float az, float bx, float by, float bz) {
    return MortonCode6D()(make_float3(ax, ay, az), make_float3(bx, by, bz));
}
// Host-side batch form for the exhaustive / large-sample tests.
void ref_morton3d_batch(const float* p3, int n, uint32_t* out) {
    for (int i = 0; i < n; ++i) out[i] = MortonCode3D()(make_float3(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2]));
}
}
