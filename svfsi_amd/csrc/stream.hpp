// stream.hpp -- CSR-stream block products for dof 3 and 4.
//
// The rows of the block-CSR matrix are cut into tiles, each a run of whole
// rows holding at most T blocks (built once per pattern on the host).  One
// 256-lane workgroup takes a tile: every lane loads T/256 of the tile's
// blocks, all loads issued before the first use, and stores the DOF
// products K_b x(:, col_b) in LDS; then S lanes per row sum the row's
// products and shuffle the S partial sums together.  In the load phase every
// lane is busy whatever the row lengths, which the row-per-lane-group
// layouts lose on meshes that mix short and long rows (tet10 vertex rows
// hold up to 65 blocks, edge rows about 20).
#pragma once

#include "devutil.hpp"

namespace svfsi {

constexpr int kStreamS = 8; // lanes per row in the summing phase

// the 16 values of a 4x4 block (64 / 128 bytes, 16-byte aligned)
template <bool NT>
__device__ __forceinline__ void load16(const float* __restrict__ u, float v[16])
{
    typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        f4 a;
        if constexpr (NT) a = __builtin_nontemporal_load(reinterpret_cast<const f4*>(u) + h);
        else a = *(reinterpret_cast<const f4*>(u) + h);
        v[4 * h] = a.x; v[4 * h + 1] = a.y; v[4 * h + 2] = a.z; v[4 * h + 3] = a.w;
    }
}

template <bool NT>
__device__ __forceinline__ void load16(const double* __restrict__ u, double v[16])
{
    typedef double d2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        d2 a;
        if constexpr (NT) a = __builtin_nontemporal_load(reinterpret_cast<const d2*>(u) + h);
        else a = *(reinterpret_cast<const d2*>(u) + h);
        v[2 * h] = a.x; v[2 * h + 1] = a.y;
    }
}

template <int DOF, bool NT, typename VT>
__device__ __forceinline__ void load_block(const VT* __restrict__ u, VT v[DOF * DOF])
{
    if constexpr (DOF == 4) load16<NT>(u, v);
    else load9<NT>(u, v);
}

// x(:, col): four values with vector loads (16-byte aligned rows), three as
// scalars
__device__ __forceinline__ void load_x(const float* __restrict__ u, float v[4])
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 a = *reinterpret_cast<const f4*>(u);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void load_x(const double* __restrict__ u, double v[4])
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 a = *reinterpret_cast<const d2*>(u), b = *(reinterpret_cast<const d2*>(u) + 1);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
template <int DOF, typename VT>
__device__ __forceinline__ void load_xv(const VT* __restrict__ u, VT v[DOF])
{
    if constexpr (DOF == 4) load_x(u, v);
    else {
#pragma unroll
        for (int m = 0; m < DOF; ++m) v[m] = u[m];
    }
}

// The rows of tile b: ep(row, i, v, s0, s1, s2, s3) runs in lane i < DOF of
// each row's group, v being component i of the row's K x and s0..s3 all of
// its components (s3 = 0 for dof 3); scalars rather than an array, so that
// selecting a lane's component needs no indexed register array, which the
// compiler would place in scratch memory.  With spmv false the products are
// skipped and the sums are zero.  Ends with a barrier, so that the
// workgroup may go on to another tile.
template <typename VT, int DOF, int T, bool NT, class EP>
__device__ __forceinline__ void stream_tile(int b, const int* __restrict__ tile, const int* __restrict__ rowPtr,
                                            const int* __restrict__ colPtr, const VT* __restrict__ val,
                                            const VT* __restrict__ x, bool spmv, EP&& ep)
{
    static_assert(T % 256 == 0, "T: a multiple of the workgroup size");
    static_assert(DOF == 3 || DOF == 4, "dof 3 or 4");
    constexpr int U = T / 256, S = kStreamS, DD = DOF * DOF;
    __shared__ VT sP[DOF][T];
    __shared__ int sR[T + 1];
    const int tid = threadIdx.x;
    const int r0 = tile[b], r1 = tile[b + 1];
    const int k0 = rowPtr[r0];
    const int nb = rowPtr[r1] - k0, nr = r1 - r0;
    for (int i = tid; i <= nr; i += 256) sR[i] = rowPtr[r0 + i] - k0;
    if (spmv) {
        int c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = u * 256 + tid;
            c[u] = j < nb ? colPtr[k0 + j] : 0;
        }
        VT a[U][DD];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = u * 256 + tid;
            if (j < nb) load_block<DOF, NT>(val + (size_t)(k0 + j) * DD, a[u]);
            else
#pragma unroll
                for (int q = 0; q < DD; ++q) a[u][q] = 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = u * 256 + tid;
            VT xv[DOF];
            load_xv<DOF>(x + (size_t)c[u] * DOF, xv);
            if (j < nb) {
#pragma unroll
                for (int i = 0; i < DOF; ++i) {
                    VT p = a[u][i * DOF] * xv[0];
#pragma unroll
                    for (int m = 1; m < DOF; ++m) p += a[u][i * DOF + m] * xv[m];
                    sP[i][j] = p;
                }
            }
        }
    }
    __syncthreads();
    const int g = tid / S, l = tid % S;
    for (int rr = g; rr < nr; rr += 256 / S) {
        VT s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        if (spmv) {
            const int ke = sR[rr + 1];
            for (int j = sR[rr] + l; j < ke; j += S) {
                s0 += sP[0][j];
                s1 += sP[1][j];
                s2 += sP[2][j];
                if constexpr (DOF == 4) s3 += sP[DOF - 1][j];
            }
#pragma unroll
            for (int o = S >> 1; o > 0; o >>= 1) {
                s0 += __shfl_xor(s0, o, kWave);
                s1 += __shfl_xor(s1, o, kWave);
                s2 += __shfl_xor(s2, o, kWave);
                if constexpr (DOF == 4) s3 += __shfl_xor(s3, o, kWave);
            }
        }
        const VT v = l == 0 ? s0 : (l == 1 ? s1 : (l == 2 ? s2 : s3));
        if (l < DOF) ep(r0 + rr, l, v, s0, s1, s2, s3);
    }
    __syncthreads();
}

} // namespace svfsi
