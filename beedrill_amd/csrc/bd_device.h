// bd_device.h -- wavefront helpers shared by the search kernels (64-lane
// waves of CDNA).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bd {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Wave-wide scans and reductions with DPP lane moves (no LDS round trips):
// row_shr 1, 2, 4, 8 scan each 16-lane row, then row_bcast 15 and 31 carry
// the row totals into the rows above.  Lanes a move cannot reach read 0.
// Every lane of the wave must be active; results are inclusive (lane i
// holds the combination of lanes 0..i), reductions read lane 63.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp_move64(uint64_t x) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(x >> 32), CTRL, ROW_MASK, 0xf, true);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

template <typename Op>
__device__ __forceinline__ uint64_t wave_scan64(uint64_t x, Op op) {
    x = op(x, dpp_move64<0x111, 0xf>(x));  // row_shr:1
    x = op(x, dpp_move64<0x112, 0xf>(x));  // row_shr:2
    x = op(x, dpp_move64<0x114, 0xf>(x));  // row_shr:4
    x = op(x, dpp_move64<0x118, 0xf>(x));  // row_shr:8
    x = op(x, dpp_move64<0x142, 0xa>(x));  // row_bcast:15 into rows 1, 3
    x = op(x, dpp_move64<0x143, 0xc>(x));  // row_bcast:31 into rows 2, 3
    return x;
}

__device__ __forceinline__ uint64_t lane63(uint64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

// x from lane `src`, which must be the same in every lane (a scalar read,
// no LDS round trip).
__device__ __forceinline__ int32_t lane_bcast(int32_t x, int src) {
    return __builtin_amdgcn_readlane(x, src);
}
__device__ __forceinline__ int64_t lane_bcast(int64_t x, int src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)x, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)x >> 32), src);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int32_t wave_incl_scan32(int32_t x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, true);
    return x;
}

__device__ __forceinline__ long long wave_incl_scan(long long x) {
    return (long long)wave_scan64((uint64_t)x, [](uint64_t p, uint64_t q) { return p + q; });
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
    return lane63(wave_scan64(x, [](uint64_t p, uint64_t q) { return p + q; }));
}

__device__ __forceinline__ unsigned long long wave_max(unsigned long long x) {
    return lane63(wave_scan64(x, [](uint64_t p, uint64_t q) { return p > q ? p : q; }));
}

__device__ __forceinline__ unsigned long long wave_min(unsigned long long x) {
    return lane63(wave_scan64(x, [](uint64_t p, uint64_t q) { return p < q ? p : q; }));
}

__device__ __forceinline__ unsigned long long wave_or(unsigned long long x) {
    return lane63(wave_scan64(x, [](uint64_t p, uint64_t q) { return p | q; }));
}

// Position of the k-th set bit (k from 0) of x.
__device__ __forceinline__ int select_bit(uint64_t x, int k) {
    int pos = 0;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const uint64_t low = x & ((1ULL << s) - 1);
        const int c = __popcll(low);
        if (k >= c) {
            k -= c;
            x >>= s;
            pos += s;
        }
    }
    return pos;
}

}  // namespace bd
