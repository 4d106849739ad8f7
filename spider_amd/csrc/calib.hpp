// Measured rate peaks for the insertion's atomic roofline: the LDS atomic the
// brick kernel issues for every tap (ds_add_u64, all 64 lanes active, one
// private slot per lane so no two lanes share a bank) and the non-returning
// fp32 global atomic add its flush issues (consecutive lanes on consecutive
// words, spread over a caller-sized buffer).
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "insert.hpp"

namespace spbp {

constexpr int CAL_LDS_ITER = 4096;   // ds_add_u64 per lane per launch
constexpr int CAL_GLOBAL_ITER = 64;  // global atomic adds per lane per launch

// NT threads: wave w owns the 256 slots [256 w, 256 w + 256); lane l adds to
// slot 256 w + l + 64 (k mod 4) at step k.
template <int NT>
__global__ __launch_bounds__(NT) void cal_lds_kernel(unsigned long long* __restrict__ out) {
    __shared__ unsigned long long buf[NT * 4];
    for (int i = threadIdx.x; i < NT * 4; i += NT) buf[i] = 0ull;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int base = (threadIdx.x >> 6) * 256 + lane;
    unsigned long long v = (unsigned long long)lane + 1ull;
    for (int k = 0; k < CAL_LDS_ITER; ++k) {
        lds_add(buf, (base + ((k & 3) << 6)) * 8, v);
        v += 3ull;
    }
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = buf[0];
}

// Thread t of the grid starts at word t and advances by the grid's width
// each step, wrapping inside n words; the grid's width must not exceed n.
__global__ __launch_bounds__(TPB) void cal_global_kernel(float* __restrict__ buf, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (int k = 0; k < CAL_GLOBAL_ITER; ++k) {
        unsafeAtomicAdd(&buf[i], 1.0f);
        i += stride;
        if (i >= n) i -= n;
    }
}

}  // namespace spbp
