// Shared helpers for the gfx950 kernels and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include "../../include/bubbledynamics.h"

#define BD_HIP_CHECK(expr)                                                       \
    do {                                                                         \
        hipError_t _e = (expr);                                                  \
        if (_e != hipSuccess) {                                                  \
            fprintf(stderr, "bubbledynamics: %s failed: %s (%s:%d)\n", #expr,    \
                    hipGetErrorString(_e), __FILE__, __LINE__);                  \
            return BD_ERR_HIP;                                                   \
        }                                                                        \
    } while (0)

#define BD_LAUNCH_CHECK()                                                        \
    do {                                                                         \
        hipError_t _e = hipGetLastError();                                       \
        if (_e != hipSuccess) {                                                  \
            fprintf(stderr, "bubbledynamics: launch failed: %s (%s:%d)\n",       \
                    hipGetErrorString(_e), __FILE__, __LINE__);                  \
            return BD_ERR_HIP;                                                   \
        }                                                                        \
    } while (0)

namespace bd {

constexpr int kWave = 64;   // CDNA wavefront width
constexpr int kXcds = 8;    // MI355X: 8 XCDs, workgroups dispatched round-robin

// Workgroups are handed to the XCDs round-robin (block b -> XCD b % 8).  This
// remap gives every XCD a contiguous range of logical tile ids so that tiles
// that share halo data also share an L2.  Bijective for any nblocks.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    const int q = nblocks / kXcds, r = nblocks % kXcds;
    const int x = bid % kXcds, k = bid / kXcds;
    const int start = x * q + (x < r ? x : r);
    return start + k;
}

// Deterministic wave64 sum (fixed butterfly order).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Deterministic block sum of up to 16 waves; result valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds /* >= NT/64 */) {
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NT / kWave; ++i) s += lds[i];
    }
    __syncthreads();
    return s;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace bd
