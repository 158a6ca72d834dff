// devutil.hpp -- device helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace svfsi {

constexpr int kWave = 64;

// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD
// b % 8).  xcd_block maps the hardware block index to a logical one so that
// each XCD processes one contiguous eighth of the grid (a bijection for any
// grid size): the rows of an XCD are then contiguous, and the vector entries
// they gather mostly stay in that XCD's own L2.
__device__ __forceinline__ int xcd_block(int on)
{
    const unsigned b = blockIdx.x;
    if (!on) return (int)b;
    const unsigned nb = gridDim.x, q = nb >> 3, r = nb & 7u, x = b & 7u, j = b >> 3;
    return (int)(x * q + (x < r ? x : r) + j);
}

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

// Row i of a row-major 4x4 block times a 4-vector (dof-4 SpMV helpers).
// NT: streaming (nontemporal) loads of the matrix values, for matrices that
// do not fit the last-level cache; otherwise ordinary loads, so that a
// matrix reused across Krylov steps can stay resident in the 256 MB MALL.
template <bool NT>
__device__ __forceinline__ double dot4_nt(const double* __restrict__ v, const double* __restrict__ x)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    d2 a, b;
    if constexpr (NT) {
        a = __builtin_nontemporal_load(reinterpret_cast<const d2*>(v));
        b = __builtin_nontemporal_load(reinterpret_cast<const d2*>(v) + 1);
    } else {
        a = *reinterpret_cast<const d2*>(v);
        b = *(reinterpret_cast<const d2*>(v) + 1);
    }
    const d2 xa = *reinterpret_cast<const d2*>(x);
    const d2 xb = *(reinterpret_cast<const d2*>(x) + 1);
    return a.x * xa.x + a.y * xa.y + b.x * xb.x + b.y * xb.y;
}

// FP32 operands, FP32 sums (the inner operator of the mixed solve computes in
// its storage precision)
template <bool NT>
__device__ __forceinline__ float dot4_nt(const float* __restrict__ v, const float* __restrict__ x)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 a;
    if constexpr (NT)
        a = __builtin_nontemporal_load(reinterpret_cast<const f4*>(v));
    else
        a = *reinterpret_cast<const f4*>(v);
    const f4 xa = *reinterpret_cast<const f4*>(x);
    return a.x * xa.x + a.y * xa.y + a.z * xa.z + a.w * xa.w;
}


// One dof-3 block row of K x, L lanes per row, each lane whole 3x3 blocks
// (blocks k0 + l, k0 + l + L, ...; two in flight): the lane reads a block's
// nine contiguous values (two 16-byte and one 4-byte load in FP32, four
// 16-byte and one 8-byte in FP64) and x(:, col) once, with one column index
// per block, and accumulates the three row sums in AT.  Consecutive lanes
// read consecutive blocks, so a row's values stream contiguously.  The L
// partial sums are then combined by lane shuffles: every lane of the group
// returns the row's three sums.
template <bool NT, typename VT>
__device__ __forceinline__ VT ld_nt(const VT* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// The nine values of a 3x3 block with the fewest load instructions: the
// blocks of a row are only 4-byte (FP32) or 8-byte (FP64) aligned, and
// vector types declared with that alignment let the compiler issue
// dwordx4 loads on them (two plus one dword in FP32, four plus one dwordx2
// in FP64) instead of nine scalar loads.
typedef float f4_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef double d2_a8 __attribute__((ext_vector_type(2), aligned(8)));

template <bool NT>
__device__ __forceinline__ void load9(const float* __restrict__ u, float v[9])
{
    f4_a4 a, b;
    if constexpr (NT) {
        a = __builtin_nontemporal_load(reinterpret_cast<const f4_a4*>(u));
        b = __builtin_nontemporal_load(reinterpret_cast<const f4_a4*>(u + 4));
        v[8] = __builtin_nontemporal_load(u + 8);
    } else {
        a = *reinterpret_cast<const f4_a4*>(u);
        b = *reinterpret_cast<const f4_a4*>(u + 4);
        v[8] = u[8];
    }
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

template <bool NT>
__device__ __forceinline__ void load9(const double* __restrict__ u, double v[9])
{
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        d2_a8 a;
        if constexpr (NT) a = __builtin_nontemporal_load(reinterpret_cast<const d2_a8*>(u + 2 * h));
        else a = *reinterpret_cast<const d2_a8*>(u + 2 * h);
        v[2 * h] = a.x;
        v[2 * h + 1] = a.y;
    }
    if constexpr (NT) v[8] = __builtin_nontemporal_load(u + 8);
    else v[8] = u[8];
}

template <typename VT, typename XT, typename AT, int L, bool NT>
__device__ __forceinline__ void row3(const int* __restrict__ rowPtr, const int* __restrict__ colPtr,
                                     const VT* __restrict__ val, const XT* __restrict__ x, int row, int l,
                                     bool active, AT a[3])
{
    AT s0 = 0, s1 = 0, s2 = 0, t0 = 0, t1 = 0, t2 = 0;
    if (active) {
        const int e = rowPtr[row + 1];
        int k = rowPtr[row] + l;
        for (; k + L < e; k += 2 * L) {
            VT a[9], b[9];
            load9<NT>(val + (size_t)k * 9, a);
            load9<NT>(val + (size_t)(k + L) * 9, b);
            const XT* xu = x + (size_t)colPtr[k] * 3;
            const XT* xw = x + (size_t)colPtr[k + L] * 3;
            const AT u0 = (AT)xu[0], u1 = (AT)xu[1], u2 = (AT)xu[2];
            const AT w0 = (AT)xw[0], w1 = (AT)xw[1], w2 = (AT)xw[2];
            s0 += (AT)a[0] * u0 + (AT)a[1] * u1 + (AT)a[2] * u2;
            s1 += (AT)a[3] * u0 + (AT)a[4] * u1 + (AT)a[5] * u2;
            s2 += (AT)a[6] * u0 + (AT)a[7] * u1 + (AT)a[8] * u2;
            t0 += (AT)b[0] * w0 + (AT)b[1] * w1 + (AT)b[2] * w2;
            t1 += (AT)b[3] * w0 + (AT)b[4] * w1 + (AT)b[5] * w2;
            t2 += (AT)b[6] * w0 + (AT)b[7] * w1 + (AT)b[8] * w2;
        }
        if (k < e) {
            VT a[9];
            load9<NT>(val + (size_t)k * 9, a);
            const XT* xu = x + (size_t)colPtr[k] * 3;
            const AT u0 = (AT)xu[0], u1 = (AT)xu[1], u2 = (AT)xu[2];
            s0 += (AT)a[0] * u0 + (AT)a[1] * u1 + (AT)a[2] * u2;
            s1 += (AT)a[3] * u0 + (AT)a[4] * u1 + (AT)a[5] * u2;
            s2 += (AT)a[6] * u0 + (AT)a[7] * u1 + (AT)a[8] * u2;
        }
    }
    a[0] = s0 + t0;
    a[1] = s1 + t1;
    a[2] = s2 + t2;
#pragma unroll
    for (int o = L >> 1; o > 0; o >>= 1) {
        a[0] += __shfl_xor(a[0], o, kWave);
        a[1] += __shfl_xor(a[1], o, kWave);
        a[2] += __shfl_xor(a[2], o, kWave);
    }
}

} // namespace svfsi
