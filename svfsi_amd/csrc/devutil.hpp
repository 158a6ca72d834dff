// devutil.hpp -- device helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace svfsi {

constexpr int kWave = 64;

// Sum over the 64 lanes of a wavefront (butterfly; every lane gets the sum).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Sum over a block of BS threads; result valid in every thread.
// lds must hold BS/64 doubles.  Fixed order -> deterministic.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* lds)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < BS / kWave; ++i) s += lds[i];
    return s;
}

// Coherent (L1-bypassing) load of a double written by another workgroup of
// the same kernel, used by the last-block finalisation of reductions.
__device__ __forceinline__ double load_agent(const double* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Last-block election: returns true in every thread of the block that
// finished last.  counter must start at 0; it is reset by the winner.
__device__ __forceinline__ bool last_block_done(unsigned int* counter, int* lds_flag)
{
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int prev = atomicAdd(counter, 1u);
        *lds_flag = (prev == gridDim.x * gridDim.y - 1) ? 1 : 0;
    }
    __syncthreads();
    bool last = *lds_flag != 0;
    if (last) {
        __threadfence();
        if (threadIdx.x == 0) *counter = 0u;
    }
    return last;
}

} // namespace svfsi
