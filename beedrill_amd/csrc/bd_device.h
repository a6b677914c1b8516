// bd_device.h -- wavefront helpers shared by the search kernels (64-lane
// waves of CDNA).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bd {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += (unsigned long long)__shfl_xor((long long)x, o);
    return x;
}

__device__ __forceinline__ unsigned long long wave_max(unsigned long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long y = (unsigned long long)__shfl_xor((long long)x, o);
        x = y > x ? y : x;
    }
    return x;
}

__device__ __forceinline__ unsigned long long wave_or(unsigned long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x |= (unsigned long long)__shfl_xor((long long)x, o);
    return x;
}

__device__ __forceinline__ long long wave_incl_scan(long long x) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x;
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
