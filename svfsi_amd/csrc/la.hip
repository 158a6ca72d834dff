// la.hip -- block-CSR SpMV, reductions and diagonal scaling for gfx950.
//
// Block SpMV (FSILS_SPARMULVV, svFSILS/SPARMUL.f:51-133).  dof 3 and 4 run
// on the pattern's row tiles (k_sspmv, k_spoly: stream_tile in stream.hpp,
// one workgroup per tile of whole rows, flat block loads, row sums in LDS)
// where stream_tiles selects them; the row-per-lane-group kernels below
// serve the other cases.  Generic mapping (k_bspmv): a group
// of DOF*DOF lanes owns one block row; lane r = i*DOF + j always reads entry
// (i,j) of each block, so a group streams the row's contiguous blocks with
// one coalesced DOF*DOF-element load per block (72 B FP64 dof 3, 128 B FP64
// dof 4) and 64/(DOF*DOF) rows share a wavefront (7 rows for dof 3, 4 rows
// for dof 4).  The per-lane partial sums over j are combined with lane
// shuffles and the j == 0 lane stores KU(i,row).  Val is read with
// non-temporal loads (streamed once per product) so that the gathered
// vector U keeps the L2.
#include "internal.hpp"
#include "devutil.hpp"
#include "stream.hpp"

#include <hip/hip_ext.h>

#include <cmath>
#include <cstdlib>
#include <type_traits>

namespace svfsi {

template <int DOF, typename VT, typename XT>
__global__ __launch_bounds__(256) void k_bspmv(int nNo, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colPtr,
                                               const VT* __restrict__ val,
                                               const XT* __restrict__ x,
                                               XT* __restrict__ y,
                                               const int* __restrict__ stop)
{
    if (stop && *stop) return;
    constexpr int DD = DOF * DOF;
    constexpr int G = kWave / DD;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int g = lane / DD;
    const int r = lane - g * DD;
    const int i = r / DOF;
    const int j = r - i * DOF;
    const int row = (blockIdx.x * (blockDim.x >> 6) + wave) * G + g;
    const bool active = (g < G) && (row < nNo);
    double acc = 0.0;
    if (active) {
        int k = rowPtr[row];
        const int e = rowPtr[row + 1];
        double a0 = 0.0, a1 = 0.0;
        for (; k + 1 < e; k += 2) {
            const int c0 = colPtr[k], c1 = colPtr[k + 1];
            const double v0 = (double)__builtin_nontemporal_load(val + (size_t)k * DD + r);
            const double v1 = (double)__builtin_nontemporal_load(val + (size_t)(k + 1) * DD + r);
            a0 += v0 * (double)x[(size_t)c0 * DOF + j];
            a1 += v1 * (double)x[(size_t)c1 * DOF + j];
        }
        if (k < e) {
            const int c0 = colPtr[k];
            a0 += (double)__builtin_nontemporal_load(val + (size_t)k * DD + r) * (double)x[(size_t)c0 * DOF + j];
        }
        acc = a0 + a1;
    }
    if constexpr (DOF == 4) {
        acc += __shfl_xor(acc, 1, kWave);
        acc += __shfl_xor(acc, 2, kWave);
    } else if constexpr (DOF == 2) {
        acc += __shfl_xor(acc, 1, kWave);
    } else {
        double t = 0.0;
#pragma unroll
        for (int q = 1; q < DOF; ++q) t += __shfl_down(acc, q, kWave);
        acc += t;
    }
    if (active && j == 0) y[(size_t)row * DOF + i] = (XT)acc;
}

// Value-streaming block SpMV for dof 2 and 3, whose dof x dof blocks are
// not 16-byte multiples: LPR lanes own a block row and read its values (one
// contiguous array of nb*dof^2 entries) in coalesced strides of LPR values;
// the lane holding value v adds val[v] x(col(v/dd), (v % dd) % dof) to output
// (v % dd) / dof.  The three (or two) partial sums then meet by shuffles.
template <int DOF, typename VT, typename XT, int LPR, bool NT>
__global__ __launch_bounds__(256) void k_bspmv_vs(int nNo, const int* __restrict__ rowPtr,
                                                  const int* __restrict__ colPtr,
                                                  const VT* __restrict__ val,
                                                  const XT* __restrict__ x,
                                                  XT* __restrict__ y,
                                                  const int* __restrict__ stop,
                                                  const VT* __restrict__ dinv = nullptr, int xcd = 0)
{
    if (stop && *stop) return;
    constexpr int DD = DOF * DOF;
    const int t = xcd_block(xcd) * 256 + threadIdx.x;
    const int row = t / LPR;
    const int l = t - row * LPR;
    double acc[DOF];
#pragma unroll
    for (int i = 0; i < DOF; ++i) acc[i] = 0.0;
    if (row < nNo) {
        const int k0 = rowPtr[row];
        const int nv = (rowPtr[row + 1] - k0) * DD;
        const VT* vr = val + (size_t)k0 * DD;
        const int* cr = colPtr + k0;
        int v = l;
        for (; v + LPR < nv; v += 2 * LPR) {
            const int va = v, vb = v + LPR;
            const double a = NT ? (double)__builtin_nontemporal_load(vr + va) : (double)vr[va];
            const double b = NT ? (double)__builtin_nontemporal_load(vr + vb) : (double)vr[vb];
            const int ka = va / DD, ea = va - ka * DD, kb = vb / DD, eb = vb - kb * DD;
            const double xa = (double)x[(size_t)cr[ka] * DOF + (ea % DOF)];
            const double xb = (double)x[(size_t)cr[kb] * DOF + (eb % DOF)];
#pragma unroll
            for (int i = 0; i < DOF; ++i) {
                acc[i] += (ea / DOF == i) ? a * xa : 0.0;
                acc[i] += (eb / DOF == i) ? b * xb : 0.0;
            }
        }
        if (v < nv) {
            const double a = NT ? (double)__builtin_nontemporal_load(vr + v) : (double)vr[v];
            const int ka = v / DD, ea = v - ka * DD;
            const double xa = (double)x[(size_t)cr[ka] * DOF + (ea % DOF)];
#pragma unroll
            for (int i = 0; i < DOF; ++i) acc[i] += (ea / DOF == i) ? a * xa : 0.0;
        }
    }
#pragma unroll
    for (int i = 0; i < DOF; ++i)
#pragma unroll
        for (int o = LPR >> 1; o > 0; o >>= 1) acc[i] += __shfl_xor(acc[i], o, kWave);
    if (dinv && row < nNo) { // block-Jacobi epilogue: acc <- D^-1 acc
        double v[DOF];
#pragma unroll
        for (int i = 0; i < DOF; ++i) {
            v[i] = 0.0;
#pragma unroll
            for (int m = 0; m < DOF; ++m) v[i] += (double)dinv[(size_t)row * DD + i * DOF + m] * acc[m];
        }
#pragma unroll
        for (int i = 0; i < DOF; ++i) acc[i] = v[i];
    }
    if (row < nNo && l < DOF) {
        double out = acc[0];
#pragma unroll
        for (int i = 1; i < DOF; ++i)
            if (l == i) out = acc[i];
        y[(size_t)row * DOF + l] = (XT)out;
    }
}

// dof 3 on the block-per-lane layout: L lanes per block row, each lane
// whole 3x3 blocks (row3), sums accumulated in AT (FP32 for the FP32
// matrix, FP64 for the FP64 one).  dinv (optional): per-node inverse
// diagonal blocks, the output is then D^-1 (K x).
template <typename VT, typename XT, typename AT, int L, bool NT>
__global__ __launch_bounds__(256) void k_bspmv3(int nNo, const int* __restrict__ rowPtr,
                                                const int* __restrict__ colPtr, const VT* __restrict__ val,
                                                const XT* __restrict__ x, XT* __restrict__ y,
                                                const int* __restrict__ stop, const VT* __restrict__ dinv,
                                                int xcd)
{
    if (stop && *stop) return;
    const int t = xcd_block(xcd) * 256 + threadIdx.x;
    const int row = t / L;
    const int l = t - row * L;
    const bool active = row < nNo;
    AT a[3];
    row3<VT, XT, AT, L, NT>(rowPtr, colPtr, val, x, row, l, active, a);
    if (!active || l >= 3) return;
    AT out = l == 0 ? a[0] : (l == 1 ? a[1] : a[2]);
    if (dinv) {
        const VT* d = dinv + (size_t)row * 9 + l * 3;
        out = (AT)d[0] * a[0] + (AT)d[1] * a[1] + (AT)d[2] * a[2];
    }
    y[(size_t)row * 3 + l] = (XT)out;
}

// Lanes per block row of the block-per-lane dof-3 kernels; SVFSI_SPMV3_L =
// 4, 8, 16 or 32 overrides.
// The XCD-contiguous block order of the SpMV-shaped kernels (xcd_block),
// SVFSI_XCD_MAP = 1; off by default: measured no faster (dof-3 cube
// matrices 82.6 / 171.3 us against 83.3 / 170.8 us in hardware order, the
// 3.2M-tet cylinder's mixed step 252 against 248 ms): the gathered vector
// entries already come from L2 or the Infinity Cache.
int xcd_map()
{
    static const int v = [] {
        const char* e = getenv("SVFSI_XCD_MAP");
        return e ? (atoi(e) != 0 ? 1 : 0) : 0;
    }();
    return v;
}

int spmv3_lanes(const Lhs& L)
{
    static const int forced = [] {
        const char* e = getenv("SVFSI_SPMV3_L");
        const int v = e ? atoi(e) : 0;
        return (v == 4 || v == 8 || v == 16 || v == 32) ? v : 0;
    }();
    if (forced) return forced;
    // FP32 measurements: tet4 cube matrix (15 blocks per row) 83 us at 16
    // lanes against 108 at 8, 135 at 4 and 110 at 32; tet10 cube matrix (28
    // blocks per row) 169 us at 32 against 190 at 16 and 210 at 8
    return (size_t)L.nnz > (size_t)20 * L.nNo ? 32 : 16;
}

// Greedy row tiles: a tile takes rows while its blocks stay within T (and its
// rows within T, so that the row offsets fit the workgroup's LDS table).
static int row_tiles(const std::vector<int>& rp, int nNo, int T, std::vector<int>& out)
{
    out.clear();
    out.push_back(0);
    int r = 0;
    while (r < nNo) {
        if (rp[r + 1] - rp[r] > T) {
            out.clear();
            return 0;
        }
        int r1 = r + 1;
        while (r1 < nNo && r1 - r < T && rp[r1 + 1] - rp[r] <= T) ++r1;
        out.push_back(r1);
        r = r1;
    }
    return (int)out.size() - 1;
}

int build_row_tiles(Lhs& L)
{
    static const int Ts[3] = {256, 512, 1024};
    for (int t = 0; t < 3; ++t) {
        std::vector<int> tl;
        L.nTiles[t] = 0;
        if ((int)L.hRowPtr.size() != L.nNo + 1 || L.nNo == 0) continue;
        const int n = row_tiles(L.hRowPtr, L.nNo, Ts[t], tl);
        if (n > 0) {
            SV_TRY(L.tiles[t].ensure(tl.size() * sizeof(int)));
            SV_HIP(hipMemcpy(L.tiles[t].p, tl.data(), tl.size() * sizeof(int), hipMemcpyHostToDevice));
            L.nTiles[t] = n;
        }
    }
    return SVFSI_OK;
}

// CSR-stream products (stream_tile) measured against the row-per-lane-group
// kernels, us per SpMV (tet10 cube 40^3 / tet4 cube 80^3 for dof 3, the
// 200k / 3.2M-tet cylinders for dof 4):
//   dof 3 FP32  111 / 58 (T = 512) against 178 / 86
//   dof 3 FP64  203 / 111 (T = 512) against 331 / 182
//   dof 4 FP32  6.2 (T = 256) / 107 (T = 512) against 7.2 / 120
//   dof 4 FP64  13.1 / 224 (T = 512) against 13.2 / 239
// FP64 wants two blocks in flight per lane (T = 512: 16.2 us at T = 256 on
// the small cylinder); FP32 prefers more, smaller tiles below 2M blocks.
// FP64 dof 4 keeps k_bspmv4 below 2M blocks unless forced: it gains little
// there (13.1 against 13.2 us), and those systems keep the summation order
// their GMRESV iteration counts were pinned with (a GMRES of thousands of
// iterations moves its count by a few per cent with the rounding of the
// products); from 2M blocks on the stream kernel saves 6 % (224 / 239 us).
// SVFSI_SPMV_STREAM = 0 selects the older kernels, 256 or 512 forces T.
const int* stream_tiles(const Lhs& L, int dof, bool fp64, int& T, int& nT)
{
    static const int env = [] {
        const char* e = getenv("SVFSI_SPMV_STREAM");
        return e ? atoi(e) : -1;
    }();
    if (env == 0 || (dof != 3 && dof != 4)) return nullptr;
    if (fp64 && dof == 4 && env < 0 && (size_t)L.nnz < ((size_t)2 << 20)) return nullptr;
    int t = (fp64 || (size_t)L.nnz >= ((size_t)2 << 20)) ? 1 : 0;
    if (env == 256) t = 0;
    else if (env == 512) t = 1;
    if (!L.nTiles[t]) t ^= 1;
    if (!L.nTiles[t]) return nullptr;
    T = t ? 512 : 256;
    nT = L.nTiles[t];
    return L.tiles[t].as<int>();
}

// K x on the row tiles (optionally D^-1 K x with per-node inverse diagonal
// blocks, row-major)
template <typename VT, int DOF, int T>
__global__ __launch_bounds__(256) void k_sspmv(const int* __restrict__ tile, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colPtr, const VT* __restrict__ val,
                                               const VT* __restrict__ x, VT* __restrict__ y,
                                               const int* __restrict__ stop, const VT* __restrict__ dinv)
{
    if (stop && *stop) return;
    stream_tile<VT, DOF, T, false>(blockIdx.x, tile, rowPtr, colPtr, val, x, true,
                                   [&](int row, int i, VT v, VT s0, VT s1, VT s2, VT s3) {
                                       if (dinv) {
                                           const VT* d = dinv + (size_t)row * DOF * DOF + i * DOF;
                                           v = d[0] * s0 + d[1] * s1 + d[2] * s2;
                                           if constexpr (DOF == 4) v += d[3] * s3;
                                       }
                                       y[(size_t)row * DOF + i] = v;
                                   });
}

// dof = 4 specialisation: 4 lanes per block row, lane i owns output
// component i and streams row i of every block of the row (4 contiguous
// values: 32 B FP64 / 16 B FP32) plus the 4-vector U(:,col) with vector
// loads; no cross-lane reduction.  16 block rows per wavefront, two blocks
// in flight per lane.
// dinv (optional): per-node inverse diagonal blocks (row-major 4x4); the
// output is then D^-1 (K x) (block-Jacobi left preconditioning).
template <typename VT, int H, bool NT>
__global__ __launch_bounds__(256) void k_bspmv4(int nNo, const int* __restrict__ rowPtr,
                                                const int* __restrict__ colPtr,
                                                const VT* __restrict__ val,
                                                const VT* __restrict__ x,
                                                VT* __restrict__ y,
                                                const int* __restrict__ stop,
                                                const VT* __restrict__ dinv, int xcd)
{
    // 4*H lanes per block row: lane (h, i) takes blocks k = s + h, s + h + H,
    // ... and output component i; the H partial sums meet by lane shuffles.
    if (stop && *stop) return;
    const int t = xcd_block(xcd) * blockDim.x + threadIdx.x;
    const int row = t / (4 * H);
    const int l = t - row * 4 * H;
    const int h = l >> 2, i = l & 3;
    using AT = VT; // sums in the storage precision
    AT a0 = 0, a1 = 0;
    if (row < nNo) {
        const int e = rowPtr[row + 1];
        int k = rowPtr[row] + h;
        for (; k + H < e; k += 2 * H) {
            const int c0 = colPtr[k], c1 = colPtr[k + H];
            a0 += dot4_nt<NT>(val + (size_t)k * 16 + i * 4, x + (size_t)c0 * 4);
            a1 += dot4_nt<NT>(val + (size_t)(k + H) * 16 + i * 4, x + (size_t)c1 * 4);
        }
        if (k < e) a0 += dot4_nt<NT>(val + (size_t)k * 16 + i * 4, x + (size_t)colPtr[k] * 4);
    }
    AT acc = a0 + a1;
#pragma unroll
    for (int o = 4; o < 4 * H; o <<= 1) acc += __shfl_xor(acc, o, kWave);
    if (dinv) {
        // every lane holds (K x)(row, i); gather the row's four components
        const int base = (threadIdx.x & (kWave - 1)) & ~3;
        AT kx[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) kx[m] = __shfl(acc, base + m, kWave);
        AT v = 0;
        if (row < nNo) {
            const VT* d = dinv + (size_t)row * 16 + i * 4;
#pragma unroll
            for (int m = 0; m < 4; ++m) v += d[m] * kx[m];
        }
        acc = v;
    }
    if (row < nNo && h == 0) y[(size_t)row * 4 + i] = (VT)acc;
}

// dof 3 on the dof-4 lane mapping: 4*H lanes per block row, lane (h, i)
// takes blocks s + h, s + h + H, ... and block row i < 3 (three contiguous
// values), lane i = 3 idles; x(:, col) is read as three contiguous values.  The H partial sums meet by lane
// shuffles.
template <typename VT, int H, bool NT>
__global__ __launch_bounds__(256) void k_bspmv3r(int nNo, const int* __restrict__ rowPtr,
                                                 const int* __restrict__ colPtr, const VT* __restrict__ val,
                                                 const VT* __restrict__ x, VT* __restrict__ y,
                                                 const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int row = t / (4 * H);
    const int l = t - row * 4 * H;
    const int h = l >> 2, i = l & 3;
    VT a0 = 0, a1 = 0;
    if (row < nNo && i < 3) {
        const int e = rowPtr[row + 1];
        int k = rowPtr[row] + h;
        for (; k + H < e; k += 2 * H) {
            const int c0 = colPtr[k], c1 = colPtr[k + H];
            const VT* ua = val + (size_t)k * 9 + i * 3;
            const VT* ub = val + (size_t)(k + H) * 9 + i * 3;
            const VT* xa = x + (size_t)c0 * 3;
            const VT* xb = x + (size_t)c1 * 3;
            a0 += ld_nt<NT>(ua) * xa[0] + ld_nt<NT>(ua + 1) * xa[1] + ld_nt<NT>(ua + 2) * xa[2];
            a1 += ld_nt<NT>(ub) * xb[0] + ld_nt<NT>(ub + 1) * xb[1] + ld_nt<NT>(ub + 2) * xb[2];
        }
        if (k < e) {
            const VT* ua = val + (size_t)k * 9 + i * 3;
            const VT* xa = x + (size_t)colPtr[k] * 3;
            a0 += ld_nt<NT>(ua) * xa[0] + ld_nt<NT>(ua + 1) * xa[1] + ld_nt<NT>(ua + 2) * xa[2];
        }
    }
    VT acc = a0 + a1;
#pragma unroll
    for (int o = 4; o < 4 * H; o <<= 1) acc += __shfl_xor(acc, o, kWave);
    if (row < nNo && h == 0 && i < 3) y[(size_t)row * 3 + i] = acc;
}

// ---------------------------------------------------------------------------
// Polynomial preconditioner steps (dof 4; the block products summed in the
// inner precision, the factor's vector update in FP64 registers).
// p(B) v with 1 - a p(a) = prod_s f_s(a) is accumulated factor by factor:
// prod <- f_s(B) prod, y <- y + (the matching part of p).  For a quadratic
// factor 1 - c1 a + c2 a^2:
//   mode 1: w = B prod;           y += c1 prod - c2 w
//   mode 2: prod' = prod - c1 w + c2 B w     (the SpMV is skipped if c2 = 0)
// For a linear factor 1 - c1 a:
//   mode 0: prod' = prod - c1 B prod;  y += c1 prod
// B = D^-1 K when dinv is given (block-Jacobi), else K.  coef is a device
// array (c1, c2 per factor), so the launches do not depend on the values.
// ---------------------------------------------------------------------------
template <typename IT, int MODE, bool NT, int H = 4>
__global__ __launch_bounds__(256) void k_poly4(int nNo, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colPtr, const IT* __restrict__ val,
                                               const IT* __restrict__ dinv, const IT* __restrict__ x,
                                               const IT* __restrict__ prodOld, IT* __restrict__ w,
                                               IT* __restrict__ y, IT* __restrict__ out,
                                               const double* __restrict__ coef, int step,
                                               const int* __restrict__ stop, int xcd)
{
    if (stop && *stop) return;
    const double c1 = coef[2 * step], c2 = coef[2 * step + 1];
    const int t = xcd_block(xcd) * blockDim.x + threadIdx.x;
    const int row = t / (4 * H);
    const int l = t - row * 4 * H;
    const int h = l >> 2, i = l & 3;
    const bool skipSpmv = (MODE == 2) && c2 == 0.0;
    IT a0 = 0, a1 = 0; // sums in the inner precision
    if (row < nNo && !skipSpmv) {
        const int e = rowPtr[row + 1];
        int k = rowPtr[row] + h;
        for (; k + H < e; k += 2 * H) {
            const int cA = colPtr[k], cB = colPtr[k + H];
            a0 += dot4_nt<NT>(val + (size_t)k * 16 + i * 4, x + (size_t)cA * 4);
            a1 += dot4_nt<NT>(val + (size_t)(k + H) * 16 + i * 4, x + (size_t)cB * 4);
        }
        if (k < e) a0 += dot4_nt<NT>(val + (size_t)k * 16 + i * 4, x + (size_t)colPtr[k] * 4);
    }
    IT acc = a0 + a1;
#pragma unroll
    for (int o = 4; o < 4 * H; o <<= 1) acc += __shfl_xor(acc, o, kWave);
    if (dinv && !skipSpmv) {
        const int base = (threadIdx.x & (kWave - 1)) & ~3;
        IT kx[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) kx[m] = __shfl(acc, base + m, kWave);
        IT v = 0;
        if (row < nNo) {
            const IT* d = dinv + (size_t)row * 16 + i * 4;
#pragma unroll
            for (int m = 0; m < 4; ++m) v += d[m] * kx[m];
        }
        acc = v;
    }
    if (row >= nNo || h != 0) return;
    const size_t q = (size_t)row * 4 + i;
    if constexpr (MODE == 0) {
        const double pv = (double)x[q];
        y[q] = (IT)((double)y[q] + c1 * pv);
        out[q] = (IT)(pv - c1 * acc);
    } else if constexpr (MODE == 1) {
        const double pv = (double)x[q];
        w[q] = (IT)acc;
        y[q] = (IT)((double)y[q] + c1 * pv - c2 * acc);
    } else {
        out[q] = (IT)((double)prodOld[q] - c1 * (double)x[q] + c2 * acc);
    }
}

// dof 3: the same factor steps on the block-per-lane layout (row3, L lanes
// per block row, sums in the inner precision IT; D^-1 applied after the
// reduction).
template <typename IT, int MODE, bool NT, int L>
__global__ __launch_bounds__(256) void k_poly3(int nNo, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colPtr, const IT* __restrict__ val,
                                               const IT* __restrict__ dinv, const IT* __restrict__ x,
                                               const IT* __restrict__ prodOld, IT* __restrict__ w,
                                               IT* __restrict__ y, IT* __restrict__ out,
                                               const double* __restrict__ coef, int step,
                                               const int* __restrict__ stop, int xcd)
{
    if (stop && *stop) return;
    const double c1 = coef[2 * step], c2 = coef[2 * step + 1];
    const int t = xcd_block(xcd) * 256 + threadIdx.x;
    const int row = t / L;
    const int l = t - row * L;
    const bool skipSpmv = (MODE == 2) && c2 == 0.0;
    IT acc[3];
    row3<IT, IT, IT, L, NT>(rowPtr, colPtr, val, x, row, l, row < nNo && !skipSpmv, acc);
    if (row >= nNo || l >= 3) return;
    IT a = l == 0 ? acc[0] : (l == 1 ? acc[1] : acc[2]);
    if (dinv && !skipSpmv) {
        const IT* d = dinv + (size_t)row * 9 + l * 3;
        a = d[0] * acc[0] + d[1] * acc[1] + d[2] * acc[2];
    }
    const size_t q = (size_t)row * 3 + l;
    if constexpr (MODE == 0) {
        const double pv = (double)x[q];
        y[q] = (IT)((double)y[q] + c1 * pv);
        out[q] = (IT)(pv - c1 * a);
    } else if constexpr (MODE == 1) {
        const double pv = (double)x[q];
        w[q] = a;
        y[q] = (IT)((double)y[q] + c1 * pv - c2 * a);
    } else {
        out[q] = (IT)((double)prodOld[q] - c1 * (double)x[q] + c2 * a);
    }
}

// The factor steps of k_poly4 / k_poly3 on the row tiles (stream_tile),
// sums in the inner precision IT.
template <typename IT, int DOF, int MODE, int T>
__global__ __launch_bounds__(256) void k_spoly(const int* __restrict__ tile, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colPtr, const IT* __restrict__ val,
                                               const IT* __restrict__ dinv, const IT* __restrict__ x,
                                               const IT* __restrict__ prodOld, IT* __restrict__ w,
                                               IT* __restrict__ y, IT* __restrict__ out,
                                               const double* __restrict__ coef, int step,
                                               const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const double c1 = coef[2 * step], c2 = coef[2 * step + 1];
    const bool skipSpmv = (MODE == 2) && c2 == 0.0;
    stream_tile<IT, DOF, T, false>(
        blockIdx.x, tile, rowPtr, colPtr, val, x, !skipSpmv, [&](int row, int i, IT a, IT s0, IT s1, IT s2, IT s3) {
            if (dinv && !skipSpmv) {
                const IT* d = dinv + (size_t)row * DOF * DOF + i * DOF;
                a = d[0] * s0 + d[1] * s1 + d[2] * s2;
                if constexpr (DOF == 4) a += d[3] * s3;
            }
            const size_t q = (size_t)row * DOF + i;
            if constexpr (MODE == 0) {
                const double pv = (double)x[q];
                y[q] = (IT)((double)y[q] + c1 * pv);
                out[q] = (IT)(pv - c1 * a);
            } else if constexpr (MODE == 1) {
                const double pv = (double)x[q];
                w[q] = a;
                y[q] = (IT)((double)y[q] + c1 * pv - c2 * a);
            } else {
                out[q] = (IT)((double)prodOld[q] - c1 * (double)x[q] + c2 * a);
            }
        });
}

// y += c1 prod (the last, linear factor: no product needed)
template <typename IT>
__global__ __launch_bounds__(256) void k_poly_last(size_t n, const IT* __restrict__ prod, IT* __restrict__ y,
                                                   const double* __restrict__ coef, int step,
                                                   const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const double c1 = coef[2 * step];
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        y[q] = (IT)((double)y[q] + c1 * (double)prod[q]);
}

template <typename IT>
int launch_poly(const Lhs& L, int dof, int mode, const IT* K, const IT* dinv, const IT* x,
                const IT* prodOld, IT* w, IT* y, IT* out, const double* coef, int step, hipStream_t s,
                const int* stop)
{
    const int xm = xcd_map();
    if (dof != 3 && dof != 4) {
        set_error("polynomial preconditioner: dof %d not supported (3, 4)", dof);
        return SVFSI_ERR_NOTSUP;
    }
    const int dd = dof * dof;
    const bool nt = (size_t)L.nnz * dd * sizeof(IT) > ((size_t)128 << 20);
    const int* rp = L.rowPtr.as<int>();
    const int* cp = L.colPtr.as<int>();
    // algorithmic bytes: matrix, indices, x, the vectors the mode touches
    const double sz = (double)sizeof(IT);
    const double base = (double)L.nnz * dd * sz + 4.0 * L.nnz + 4.0 * (L.nNo + 1) + (dinv ? L.nNo * dd * sz : 0.0);
    const double vec = (double)L.nNo * dof * sz;
    ProfScope ps(std::is_same<IT, float>::value ? 1 : 0, base + vec * (mode == 0 ? 4 : mode == 1 ? 4 : 3));
    int sT = 0, nT = 0;
    if (const int* tl = stream_tiles(L, dof, std::is_same<IT, double>::value, sT, nT)) {
#define SV_SPOLY(D, M, TT)                                                                                     \
    SV_LAUNCH(ps, (k_spoly<IT, D, M, TT>), dim3((unsigned)nT), dim3(256), 0, s, tl, rp, cp, K, dinv, x, prodOld, w, y, \
              out, coef, step, stop)
#define SV_SPOLYT(D, M)                                                                                        \
    do {                                                                                                       \
        if (sT == 512) SV_SPOLY(D, M, 512);                                                                    \
        else SV_SPOLY(D, M, 256);                                                                              \
    } while (0)
        if (dof == 4) {
            if (mode == 0) SV_SPOLYT(4, 0);
            else if (mode == 1) SV_SPOLYT(4, 1);
            else SV_SPOLYT(4, 2);
        } else {
            if (mode == 0) SV_SPOLYT(3, 0);
            else if (mode == 1) SV_SPOLYT(3, 1);
            else SV_SPOLYT(3, 2);
        }
#undef SV_SPOLYT
#undef SV_SPOLY
        SV_HIP(hipGetLastError());
        return SVFSI_OK;
    }
    const int l3 = spmv3_lanes(L);
    const unsigned g3 = (unsigned)(((size_t)L.nNo * l3 + 255) / 256);
#define SV_POLY3L(M, N, LL)                                                                                    \
    SV_LAUNCH(ps, (k_poly3<IT, M, N, LL>), dim3(g3), dim3(256), 0, s, L.nNo, rp, cp, K, dinv, x, prodOld, w, y, out, \
              coef, step, stop, xm)
#define SV_POLY3(M, N)                                                                                         \
    do {                                                                                                       \
        if (l3 == 4) SV_POLY3L(M, N, 4);                                                                       \
        else if (l3 == 16) SV_POLY3L(M, N, 16);                                                                \
        else if (l3 == 32) SV_POLY3L(M, N, 32);                                                                \
        else SV_POLY3L(M, N, 8);                                                                               \
    } while (0)
    // lanes per output component of the dof-4 kernel (tuning knob
    // SVFSI_POLY_H = 2, 4 or 8)
    static const int ph = [] {
        const char* e = getenv("SVFSI_POLY_H");
        const int v = e ? atoi(e) : 4;
        return (v == 2 || v == 8) ? v : 4;
    }();
#define SV_POLYH(M, N, HH)                                                                                       \
    SV_LAUNCH(ps, (k_poly4<IT, M, N, HH>), dim3((unsigned)(((size_t)L.nNo * 4 * HH + 255) / 256)), dim3(256), 0, s, \
              L.nNo, rp, cp, K, dinv, x, prodOld, w, y, out, coef, step, stop, xm)
#define SV_POLY4(M, N)                                                                                           \
    do {                                                                                                         \
        if (ph == 2) SV_POLYH(M, N, 2);                                                                          \
        else if (ph == 8) SV_POLYH(M, N, 8);                                                                     \
        else SV_POLYH(M, N, 4);                                                                                  \
    } while (0)
    if (dof == 4) {
        if (mode == 0) { if (nt) SV_POLY4(0, true); else SV_POLY4(0, false); }
        else if (mode == 1) { if (nt) SV_POLY4(1, true); else SV_POLY4(1, false); }
        else { if (nt) SV_POLY4(2, true); else SV_POLY4(2, false); }
#undef SV_POLY4
#undef SV_POLYH
    } else {
        if (mode == 0) { if (nt) SV_POLY3(0, true); else SV_POLY3(0, false); }
        else if (mode == 1) { if (nt) SV_POLY3(1, true); else SV_POLY3(1, false); }
        else { if (nt) SV_POLY3(2, true); else SV_POLY3(2, false); }
    }
#undef SV_POLY3
#undef SV_POLY3L
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

template <typename IT>
int launch_poly_last(size_t n, const IT* prod, IT* y, const double* coef, int step, hipStream_t s,
                     const int* stop)
{
    size_t g = (n + 255) / 256;
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL(k_poly_last<IT>, dim3((unsigned)g), dim3(256), 0, s, n, prod, y, coef, step, stop);
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}
template int launch_poly<float>(const Lhs&, int, int, const float*, const float*, const float*, const float*, float*,
                                float*, float*, const double*, int, hipStream_t, const int*);
template int launch_poly<double>(const Lhs&, int, int, const double*, const double*, const double*, const double*,
                                 double*, double*, double*, const double*, int, hipStream_t, const int*);
template int launch_poly_last<float>(size_t, const float*, float*, const double*, int, hipStream_t, const int*);
template int launch_poly_last<double>(size_t, const double*, double*, const double*, int, hipStream_t, const int*);

// Inverse of each node's diagonal block (dof 4, Gauss-Jordan with partial
// pivoting in FP64), stored FP32 row-major.  Rows and columns zeroed by the
// Dirichlet mask get a unit diagonal first.
constexpr double kBlockCond = 1e6;

template <int DOF, typename OT, typename VT = double>
__global__ __launch_bounds__(256) void k_blockdiag_inv(int nNo, const int* __restrict__ diagPtr,
                                                       const VT* __restrict__ val, OT* __restrict__ dinv)
{
    constexpr int DD = DOF * DOF;
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nNo) return;
    double a[DOF][2 * DOF];
    const VT* b = val + (size_t)diagPtr[r] * DD;
#pragma unroll
    for (int p = 0; p < DOF; ++p)
#pragma unroll
        for (int q = 0; q < DOF; ++q) {
            a[p][q] = b[p * DOF + q];
            a[p][DOF + q] = (p == q) ? 1.0 : 0.0;
        }
#pragma unroll
    for (int p = 0; p < DOF; ++p) {
        bool zero = true;
#pragma unroll
        for (int q = 0; q < DOF; ++q) zero = zero && a[p][q] == 0.0 && a[q][p] == 0.0;
        if (zero) a[p][p] = 1.0;
    }
    double amax = 0.0;
#pragma unroll
    for (int p = 0; p < DOF; ++p)
#pragma unroll
        for (int q = 0; q < DOF; ++q) amax = fmax(amax, fabs(a[p][q]));
    bool ok = true;
#pragma unroll
    for (int c = 0; c < DOF; ++c) {
        int piv = c;
        double best = fabs(a[c][c]);
#pragma unroll
        for (int p = c + 1; p < DOF; ++p)
            if (fabs(a[p][c]) > best) { best = fabs(a[p][c]); piv = p; }
        if (best == 0.0) { ok = false; break; }
        // row swap with compile-time row indices (a lane-dependent index
        // would move the matrix to scratch memory)
#pragma unroll
        for (int p = c + 1; p < DOF; ++p)
            if (p == piv)
#pragma unroll
                for (int q = 0; q < 2 * DOF; ++q) { const double t = a[c][q]; a[c][q] = a[p][q]; a[p][q] = t; }
        const double inv = 1.0 / a[c][c];
#pragma unroll
        for (int q = 0; q < 2 * DOF; ++q) a[c][q] *= inv;
#pragma unroll
        for (int p = 0; p < DOF; ++p)
            if (p != c) {
                const double f = a[p][c];
#pragma unroll
                for (int q = 0; q < 2 * DOF; ++q) a[p][q] -= f * a[c][q];
            }
    }
    // A block whose inverse is ill-conditioned (max|B| max|B^-1| above
    // kBlockCond: a saddle-point node block with a zero pressure diagonal
    // and weak coupling) would turn the inner operator's rounding into
    // noise; such a node keeps the inverse of its diagonal instead.
    double imax = 0.0;
#pragma unroll
    for (int p = 0; p < DOF; ++p)
#pragma unroll
        for (int q = 0; q < DOF; ++q) imax = fmax(imax, fabs(a[p][DOF + q]));
    ok = ok && amax * imax <= kBlockCond;
    OT* o = dinv + (size_t)r * DD;
#pragma unroll
    for (int p = 0; p < DOF; ++p)
#pragma unroll
        for (int q = 0; q < DOF; ++q) {
            if (ok) {
                o[p * DOF + q] = (OT)a[p][DOF + q];
            } else {
                const double dpp = (double)b[p * DOF + p];
                o[p * DOF + q] = (OT)(p == q ? (dpp != 0.0 ? 1.0 / dpp : 1.0) : 0.0);
            }
        }
}

// B = D^-1 K in FP32 (block Jacobi folded into the inner matrix): block k of
// block row i becomes D_i^-1 K_k, formed in FP64 from the FP64 scaled matrix.
// DD lanes per block, one output entry each.
template <int DOF, typename OT>
__global__ __launch_bounds__(256) void k_bjscale(size_t nnz, const int* __restrict__ blockRow,
                                                 const double* __restrict__ K, const OT* __restrict__ dinv,
                                                 OT* __restrict__ B)
{
    constexpr int DD = DOF * DOF;
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t k = t / DD;
    if (k >= nnz) return;
    const int e = (int)(t - k * DD);
    const int r = e / DOF, c = e - (e / DOF) * DOF;
    const OT* d = dinv + (size_t)blockRow[k] * DD + r * DOF;
    const double* kb = K + k * DD;
    double v = 0.0;
#pragma unroll
    for (int m = 0; m < DOF; ++m) v += (double)d[m] * kb[m * DOF + c];
    B[k * DD + e] = (OT)v;
}

// x = D^-1 x per node (FP32, in place).
template <int DOF, typename IT>
__global__ __launch_bounds__(256) void k_blockdiag_apply(int nNo, const IT* __restrict__ dinv,
                                                         IT* __restrict__ x, const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nNo) return;
    IT xi[DOF];
#pragma unroll
    for (int m = 0; m < DOF; ++m) xi[m] = x[(size_t)r * DOF + m];
    const IT* d = dinv + (size_t)r * DOF * DOF;
#pragma unroll
    for (int p = 0; p < DOF; ++p) {
        double v = 0.0;
#pragma unroll
        for (int m = 0; m < DOF; ++m) v += (double)d[p * DOF + m] * (double)xi[m];
        x[(size_t)r * DOF + p] = (IT)v;
    }
}

// Seeded pseudo-random start vector in (-0.5, 0.5), zero where W = 0 (the
// Dirichlet-masked unknowns, whose rows and columns are zero).
template <typename IT>
__global__ __launch_bounds__(256) void k_rand_masked(size_t n, unsigned seed, const double* __restrict__ W,
                                                     IT* __restrict__ v)
{
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        unsigned x = (unsigned)q * 2654435761u ^ seed;
        x ^= x >> 16;
        x *= 0x7feb352du;
        x ^= x >> 15;
        x *= 0x846ca68bu;
        x ^= x >> 16;
        const float r = (float)(x >> 8) * (1.0f / 16777216.0f) - 0.5f;
        v[q] = (W[q] != 0.0) ? (IT)r : (IT)0;
    }
}

template <typename IT>
int launch_rand_masked(size_t n, unsigned seed, const double* W, IT* v, hipStream_t s)
{
    size_t g = (n + 255) / 256;
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL(k_rand_masked<IT>, dim3((unsigned)g), dim3(256), 0, s, n, seed, W, v);
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

template <typename OT>
int launch_blockdiag_inv(const Lhs& L, int dof, const double* val, OT* dinv, hipStream_t s)
{
    const dim3 g((L.nNo + 255) / 256);
    if (dof == 4)
        hipLaunchKernelGGL((k_blockdiag_inv<4, OT>), g, dim3(256), 0, s, L.nNo, L.diagPtr.as<int>(), val, dinv);
    else if (dof == 3)
        hipLaunchKernelGGL((k_blockdiag_inv<3, OT>), g, dim3(256), 0, s, L.nNo, L.diagPtr.as<int>(), val, dinv);
    else {
        set_error("block-Jacobi: dof %d not supported (3, 4)", dof);
        return SVFSI_ERR_NOTSUP;
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

// number of unmasked dofs (W != 0) whose diagonal entry is zero: rows of a
// saddle-point system without pressure stabilisation
__global__ __launch_bounds__(256) void k_count_zero_diag(int nNo, int dof, const int* __restrict__ diagPtr,
                                                         const double* __restrict__ val,
                                                         const double* __restrict__ W, int* __restrict__ count)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nNo * dof) return;
    const int i = t / dof, c = t - (t / dof) * dof;
    if (W[t] != 0.0 && val[(size_t)diagPtr[i] * dof * dof + c * dof + c] == 0.0) atomicAdd(count, 1);
}

int count_zero_diag(const Lhs& L, int dof, const double* val, const double* W, int* dCount, int& hCount,
                    hipStream_t s)
{
    hCount = 0;
    if (L.nNo == 0) return SVFSI_OK;
    SV_HIP(hipMemsetAsync(dCount, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_count_zero_diag, dim3((unsigned)(((size_t)L.nNo * dof + 255) / 256)), dim3(256), 0, s,
                       L.nNo, dof, L.diagPtr.as<int>(), val, W, dCount);
    SV_HIP(hipGetLastError());
    SV_HIP(hipMemcpyAsync(&hCount, dCount, sizeof(int), hipMemcpyDeviceToHost, s));
    SV_HIP(hipStreamSynchronize(s));
    return SVFSI_OK;
}

int launch_blockdiag_inv_f32(const Lhs& L, int dof, const float* val, float* dinv, hipStream_t s)
{
    const dim3 g((L.nNo + 255) / 256);
    if (dof == 4)
        hipLaunchKernelGGL((k_blockdiag_inv<4, float, float>), g, dim3(256), 0, s, L.nNo, L.diagPtr.as<int>(), val,
                           dinv);
    else if (dof == 3)
        hipLaunchKernelGGL((k_blockdiag_inv<3, float, float>), g, dim3(256), 0, s, L.nNo, L.diagPtr.as<int>(), val,
                           dinv);
    else {
        set_error("block-Jacobi: dof %d not supported (3, 4)", dof);
        return SVFSI_ERR_NOTSUP;
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

template <typename OT>
int launch_bjscale(const Lhs& L, int dof, const double* K, const OT* dinv, OT* B, hipStream_t s)
{
    const size_t nnz = (size_t)L.nnz;
    const dim3 g((unsigned)((nnz * dof * dof + 255) / 256));
    if (dof == 4)
        hipLaunchKernelGGL((k_bjscale<4, OT>), g, dim3(256), 0, s, nnz, L.blockRow.as<int>(), K, dinv, B);
    else if (dof == 3)
        hipLaunchKernelGGL((k_bjscale<3, OT>), g, dim3(256), 0, s, nnz, L.blockRow.as<int>(), K, dinv, B);
    else {
        set_error("block-Jacobi: dof %d not supported (3, 4)", dof);
        return SVFSI_ERR_NOTSUP;
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

template <typename IT>
int launch_blockdiag_apply(const Lhs& L, int dof, const IT* dinv, IT* x, hipStream_t s, const int* stop)
{
    const dim3 g((L.nNo + 255) / 256);
    if (dof == 4)
        hipLaunchKernelGGL((k_blockdiag_apply<4, IT>), g, dim3(256), 0, s, L.nNo, dinv, x, stop);
    else if (dof == 3)
        hipLaunchKernelGGL((k_blockdiag_apply<3, IT>), g, dim3(256), 0, s, L.nNo, dinv, x, stop);
    else {
        set_error("block-Jacobi: dof %d not supported (3, 4)", dof);
        return SVFSI_ERR_NOTSUP;
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}
template int launch_rand_masked<float>(size_t, unsigned, const double*, float*, hipStream_t);
template int launch_rand_masked<double>(size_t, unsigned, const double*, double*, hipStream_t);
template int launch_blockdiag_inv<float>(const Lhs&, int, const double*, float*, hipStream_t);
template int launch_blockdiag_inv<double>(const Lhs&, int, const double*, double*, hipStream_t);
template int launch_bjscale<float>(const Lhs&, int, const double*, const float*, float*, hipStream_t);
template int launch_bjscale<double>(const Lhs&, int, const double*, const double*, double*, hipStream_t);
template int launch_blockdiag_apply<float>(const Lhs&, int, const float*, float*, hipStream_t, const int*);
template int launch_blockdiag_apply<double>(const Lhs&, int, const double*, double*, hipStream_t, const int*);

// Transpose product: row i of K^T gathers block (col,i) through the
// transpose-position map trans[k] (pattern is structurally symmetric, as the
// node-adjacency pattern of LHSA.f always is) and reads it transposed.
template <int DOF>
__global__ __launch_bounds__(256) void k_bspmv_t(int nNo, const int* __restrict__ rowPtr,
                                                 const int* __restrict__ colPtr,
                                                 const int* __restrict__ trans,
                                                 const double* __restrict__ val,
                                                 const double* __restrict__ x,
                                                 double* __restrict__ y,
                                                 const int* __restrict__ stop)
{
    if (stop && *stop) return;
    constexpr int DD = DOF * DOF;
    constexpr int G = kWave / DD;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int g = lane / DD;
    const int r = lane - g * DD;
    const int i = r / DOF;
    const int j = r - i * DOF;
    const int row = (blockIdx.x * (blockDim.x >> 6) + wave) * G + g;
    const bool active = (g < G) && (row < nNo);
    double acc = 0.0;
    if (active) {
        for (int k = rowPtr[row]; k < rowPtr[row + 1]; ++k) {
            const int c = colPtr[k];
            acc += val[(size_t)trans[k] * DD + j * DOF + i] * x[(size_t)c * DOF + j];
        }
    }
    if constexpr (DOF == 4) {
        acc += __shfl_xor(acc, 1, kWave);
        acc += __shfl_xor(acc, 2, kWave);
    } else if constexpr (DOF == 2) {
        acc += __shfl_xor(acc, 1, kWave);
    } else {
        double t = 0.0;
#pragma unroll
        for (int q = 1; q < DOF; ++q) t += __shfl_down(acc, q, kWave);
        acc += t;
    }
    if (active && j == 0) y[(size_t)row * DOF + i] = acc;
}

static inline int spmv_grid(int nNo, int dof)
{
    const int rowsPerBlock = 4 * (kWave / (dof * dof));
    return (nNo + rowsPerBlock - 1) / rowsPerBlock;
}

template <typename VT, typename XT>
static int spmv_dispatch(const Lhs& L, int dof, const VT* K, const XT* U, XT* KU,
                         hipStream_t s, int profKernel, const int* stop, const VT* dinv = nullptr)
{
    if (L.nNo == 0) return SVFSI_OK;
    // algorithmic bytes: values, column indices, row pointers, x once, y once
    // (+ the inverse diagonal blocks of a block-Jacobi epilogue)
    ProfScope ps(profKernel, (double)L.nnz * dof * dof * sizeof(VT) + 4.0 * L.nnz + 4.0 * (L.nNo + 1) +
                                 2.0 * L.nNo * dof * sizeof(XT) +
                                 (dinv ? (double)L.nNo * dof * dof * sizeof(VT) : 0.0));
    const int* rp = L.rowPtr.as<int>();
    const int* cp = L.colPtr.as<int>();
    const int xm = xcd_map();
    if constexpr (std::is_same<VT, XT>::value) {
        int sT = 0, nT = 0;
        const int* tl = stream_tiles(L, dof, std::is_same<VT, double>::value, sT, nT);
        if (tl) {
#define SV_SSPMV(D, TT)                                                                                        \
    SV_LAUNCH(ps, (k_sspmv<VT, D, TT>), dim3((unsigned)nT), dim3(256), 0, s, tl, rp, cp, K, U, KU, stop, dinv)
            if (dof == 4) { if (sT == 512) SV_SSPMV(4, 512); else SV_SSPMV(4, 256); }
            else { if (sT == 512) SV_SSPMV(3, 512); else SV_SSPMV(3, 256); }
#undef SV_SSPMV
            SV_HIP(hipGetLastError());
            return SVFSI_OK;
        }
        if (dof == 4) {
            // lanes per output component (tuning knob SVFSI_SPMV_H = 1, 2 or 4;
            // 4 measured fastest on the 200k-tet cylinder)
            static const int H = [] {
                const char* e = getenv("SVFSI_SPMV_H");
                const int v = e ? atoi(e) : 4;
                return (v == 1 || v == 2) ? v : 4;
            }();
            // streaming loads only when the values exceed half the MALL
            // (tuning knob SVFSI_SPMV_NT = 0 / 1 overrides)
            static const int ntMode = [] {
                const char* e = getenv("SVFSI_SPMV_NT");
                return e ? atoi(e) : -1;
            }();
            const bool nt = ntMode >= 0 ? ntMode != 0 : (size_t)L.nnz * 16 * sizeof(VT) > ((size_t)128 << 20);
            const int g4 = (int)(((size_t)L.nNo * 4 * H + 255) / 256);
            if (H == 1 && nt)
                SV_LAUNCH(ps, (k_bspmv4<VT, 1, true>), dim3(g4), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, stop, dinv, xm);
            else if (H == 1)
                SV_LAUNCH(ps, (k_bspmv4<VT, 1, false>), dim3(g4), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, stop, dinv, xm);
            else if (H == 2 && nt)
                SV_LAUNCH(ps, (k_bspmv4<VT, 2, true>), dim3(g4), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, stop, dinv, xm);
            else if (H == 2)
                SV_LAUNCH(ps, (k_bspmv4<VT, 2, false>), dim3(g4), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, stop, dinv, xm);
            else if (nt)
                SV_LAUNCH(ps, (k_bspmv4<VT, 4, true>), dim3(g4), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, stop, dinv, xm);
            else
                SV_LAUNCH(ps, (k_bspmv4<VT, 4, false>), dim3(g4), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, stop, dinv, xm);
            SV_HIP(hipGetLastError());
            return SVFSI_OK;
        }
    }
    // dof 3 in FP64: the dof-4 lane mapping (k_bspmv3r, H = 4) measured
    // fastest: 331 us on the 40^3 tet10 cube matrix against 373 us value
    // streaming and 458 us block per lane, 182 against 187 us on the 80^3
    // tet4 one.  FP32 keeps block per lane (169 / 83 us against 169 / 99).
    // SVFSI_SPMV3_RH = 0 / 4 / 8 overrides (0: value streaming).
    static const int rhEnv = [] {
        const char* e = getenv("SVFSI_SPMV3_RH");
        if (!e) return -1;
        const int v = atoi(e);
        return (v == 4 || v == 8) ? v : 0;
    }();
    const int rh = rhEnv >= 0 ? rhEnv : (std::is_same<VT, double>::value ? 4 : 0);
    if (dof == 3 && rh && !dinv) {
        const bool nt = (size_t)L.nnz * 9 * sizeof(VT) > ((size_t)128 << 20);
        if constexpr (std::is_same<VT, XT>::value) {
            const unsigned g = (unsigned)(((size_t)L.nNo * 4 * rh + 255) / 256);
#define SV_R3(HH, N) SV_LAUNCH(ps, (k_bspmv3r<VT, HH, N>), dim3(g), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, stop)
            if (rh == 4) { if (nt) SV_R3(4, true); else SV_R3(4, false); }
            else { if (nt) SV_R3(8, true); else SV_R3(8, false); }
#undef SV_R3
            SV_HIP(hipGetLastError());
            return SVFSI_OK;
        }
    }
    // SVFSI_SPMV3_ROW = 1 puts FP64 on the block-per-lane layout as well
    static const bool row64 = [] {
        const char* e = getenv("SVFSI_SPMV3_ROW");
        return e && e[0] == '1';
    }();
    if (dof == 3 && (std::is_same<VT, float>::value || row64)) {
        // FP32: whole blocks per lane (fewer, wider loads; FP32 sums).  FP64
        // uses k_bspmv3r above: its 72-byte blocks per lane spread each load
        // instruction over 4.6 KB and measured slower (458 us at 32 lanes).
        const bool nt = (size_t)L.nnz * 9 * sizeof(VT) > ((size_t)128 << 20);
        const int l3 = spmv3_lanes(L);
        const unsigned g3 = (unsigned)(((size_t)L.nNo * l3 + 255) / 256);
#define SV_B3(LL, N)                                                                                           \
    SV_LAUNCH(ps, (k_bspmv3<VT, XT, VT, LL, N>), dim3(g3), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, stop, dinv, xm)
        if (l3 == 4) { if (nt) SV_B3(4, true); else SV_B3(4, false); }
        else if (l3 == 16) { if (nt) SV_B3(16, true); else SV_B3(16, false); }
        else if (l3 == 32) { if (nt) SV_B3(32, true); else SV_B3(32, false); }
        else { if (nt) SV_B3(8, true); else SV_B3(8, false); }
#undef SV_B3
        SV_HIP(hipGetLastError());
        return SVFSI_OK;
    }
    if (dof == 2 || dof == 3) {
        // value-streaming layout (tuning knob SVFSI_SPMV_LPR = 8 / 16 / 32;
        // 0 selects the block layout below)
        static const int lpr = [] {
            const char* e = getenv("SVFSI_SPMV_LPR");
            const int v = e ? atoi(e) : 16;
            return (v == 0 || v == 8 || v == 32) ? v : 16;
        }();
        const bool nt = (size_t)L.nnz * dof * dof * sizeof(VT) > ((size_t)128 << 20);
        if (lpr) {
            const unsigned gvs = (unsigned)(((size_t)L.nNo * lpr + 255) / 256);
#define SV_VS(D, P)                                                                                           \
    do {                                                                                                      \
        if (nt)                                                                                               \
            SV_LAUNCH(ps, (k_bspmv_vs<D, VT, XT, P, true>), dim3(gvs), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, \
                      stop, dinv, xm);                                                                        \
        else                                                                                                  \
            SV_LAUNCH(ps, (k_bspmv_vs<D, VT, XT, P, false>), dim3(gvs), dim3(256), 0, s, L.nNo, rp, cp, K, U,    \
                      KU, stop, dinv, xm);                                                                    \
    } while (0)
            if (dof == 2) {
                if (lpr == 8) SV_VS(2, 8); else if (lpr == 32) SV_VS(2, 32); else SV_VS(2, 16);
            } else {
                if (lpr == 8) SV_VS(3, 8); else if (lpr == 32) SV_VS(3, 32); else SV_VS(3, 16);
            }
#undef SV_VS
            SV_HIP(hipGetLastError());
            return SVFSI_OK;
        }
    }
    if (dinv) {
        set_error("block-Jacobi SpMV: needs the value-streaming layout for dof %d", dof);
        return SVFSI_ERR_NOTSUP;
    }
    const int grid = spmv_grid(L.nNo, dof);
    switch (dof) {
    case 2: SV_LAUNCH(ps, (k_bspmv<2, VT, XT>), dim3(grid), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, stop); break;
    case 3: SV_LAUNCH(ps, (k_bspmv<3, VT, XT>), dim3(grid), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, stop); break;
    case 4: SV_LAUNCH(ps, (k_bspmv<4, VT, XT>), dim3(grid), dim3(256), 0, s, L.nNo, rp, cp, K, U, KU, stop); break;
    default: set_error("spmv: dof %d not supported (2..4)", dof); return SVFSI_ERR_NOTSUP;
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

int launch_spmv(const Lhs& L, int dof, const double* K, const double* U, double* KU,
                hipStream_t s, const int* stop)
{
    return spmv_dispatch<double, double>(L, dof, K, U, KU, s, 0, stop);
}

int launch_spmv_f32(const Lhs& L, int dof, const float* K, const float* U, float* KU,
                    hipStream_t s, const int* stop, const float* dinv, int profSlot)
{
    if (dinv && dof != 4 && dof != 3) {
        set_error("block-Jacobi SpMV: dof %d not supported (3, 4)", dof);
        return SVFSI_ERR_NOTSUP;
    }
    return spmv_dispatch<float, float>(L, dof, K, U, KU, s, profSlot, stop, dinv);
}

int launch_spmv_t(const Lhs& L, int dof, const double* K, const double* U, double* KU,
                  hipStream_t s, const int* stop)
{
    if (L.nNo == 0) return SVFSI_OK;
    if (!L.symmetric) {
        set_error("transpose product needs a structurally symmetric pattern");
        return SVFSI_ERR_ARG;
    }
    const int grid = spmv_grid(L.nNo, dof);
    const int* rp = L.rowPtr.as<int>();
    const int* cp = L.colPtr.as<int>();
    const int* tr = L.trans.as<int>();
    switch (dof) {
    case 2: hipLaunchKernelGGL(k_bspmv_t<2>, dim3(grid), dim3(256), 0, s, L.nNo, rp, cp, tr, K, U, KU, stop); break;
    case 3: hipLaunchKernelGGL(k_bspmv_t<3>, dim3(grid), dim3(256), 0, s, L.nNo, rp, cp, tr, K, U, KU, stop); break;
    case 4: hipLaunchKernelGGL(k_bspmv_t<4>, dim3(grid), dim3(256), 0, s, L.nNo, rp, cp, tr, K, U, KU, stop); break;
    default: set_error("spmv_t: dof %d not supported (2..4)", dof); return SVFSI_ERR_NOTSUP;
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

// ---------------------------------------------------------------------------
// Deterministic dot product: fixed grid, per-block partials, fixed-order sum.
// ---------------------------------------------------------------------------
constexpr int kDotBlocks = 512;

__global__ __launch_bounds__(256) void k_dot_partial(size_t n, const double* __restrict__ u,
                                                     const double* __restrict__ v,
                                                     double* __restrict__ partial)
{
    __shared__ double lds[4];
    double s = 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        s += u[q] * v[q];
    s = block_sum<256>(s, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_sum_partials(int m, const double* __restrict__ partial,
                                                      double* __restrict__ out)
{
    __shared__ double lds[4];
    double s = 0.0;
    for (int q = threadIdx.x; q < m; q += 256) s += partial[q];
    s = block_sum<256>(s, lds);
    if (threadIdx.x == 0) *out = s;
}

int launch_dot(int n, const double* U, const double* V, double* partial, double* result, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(256), 0, s, (size_t)n, U, V, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, kDotBlocks, partial, result);
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

// ---------------------------------------------------------------------------
// PRECONDDIAG (svFSILS/PRECOND.f:50-145) and PREMUL/POSMUL (372-489).
// ---------------------------------------------------------------------------
__global__ void k_diag_w(int nNo, int dof, const int* __restrict__ diagPtr,
                         const double* __restrict__ val, const double* __restrict__ mask,
                         double* __restrict__ W)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nNo * dof) return;
    const int a = q / dof, i = q - a * dof;
    double d = val[(size_t)diagPtr[a] * dof * dof + i * dof + i];
    if (d == 0.0) d = 1.0;                 // PRECOND.f:104-109
    double w = 1.0 / sqrt(fabs(d));        // PRECOND.f:111
    W[q] = w * mask[q];                    // Dirichlet faces, PRECOND.f:112-121
}

// Val(:,k) <- (W_row * Val) * W_col, the PREMUL then POSMUL rounding order;
// optionally also writes an FP32 copy of the scaled matrix.
template <int DOF>
__global__ __launch_bounds__(256) void k_scale_val(size_t nEnt, const int* __restrict__ blockRow,
                                                   const int* __restrict__ colPtr,
                                                   const double* __restrict__ W,
                                                   double* __restrict__ val,
                                                   float* __restrict__ val32)
{
    constexpr int DD = DOF * DOF;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nEnt) return;
    const size_t k = t / DD;
    const int r = (int)(t - k * DD);
    const int i = r / DOF, j = r - (r / DOF) * DOF;
    const double wr = W[(size_t)blockRow[k] * DOF + i];
    const double wc = W[(size_t)colPtr[k] * DOF + j];
    const double v = (val[t] * wr) * wc;
    val[t] = v;
    if (val32) val32[t] = (float)v;
}

__global__ void k_scale_vec(size_t n, const double* __restrict__ W, double* __restrict__ R)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) R[q] = W[q] * R[q];
}

int launch_scale_vec(size_t n, const double* W, double* R, hipStream_t s)
{
    if (n == 0) return SVFSI_OK;
    hipLaunchKernelGGL(k_scale_vec, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, W, R);
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

// Dense Dirichlet multiplier mask(dof,nNo): product of face%val over the
// included Dirichlet faces (PRECOND.f:112-121, i = MIN(face%dof, dof)).
int build_mask(Lhs& L, int dof, const int* incL, hipStream_t s)
{
    std::vector<int> inc(L.faces.size(), 1);
    if (incL)
        for (size_t f = 0; f < L.faces.size(); ++f) inc[f] = incL[f] != 0;
    Work& w = L.ws;
    if (w.maskDof == dof && w.maskVersion == L.facesVersion && w.maskInc == inc) return SVFSI_OK;
    std::vector<double> m((size_t)L.nNo * dof, 1.0);
    for (size_t f = 0; f < L.faces.size(); ++f) {
        const Face& F = L.faces[f];
        if (!F.foC || !inc[f] || F.bGrp != BC_TYPE_Dir) continue;
        const int nd = F.dof < dof ? F.dof : dof;
        for (int a = 0; a < F.nNo; ++a)
            for (int i = 0; i < nd; ++i) m[(size_t)F.glob[a] * dof + i] *= F.val[(size_t)a * F.dof + i];
    }
    SV_TRY(w.mask.ensure(m.size() * sizeof(double)));
    SV_HIP(hipMemcpyAsync(w.mask.p, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice, s));
    SV_HIP(hipStreamSynchronize(s));
    w.maskDof = dof;
    w.maskVersion = L.facesVersion;
    w.maskInc = inc;
    return SVFSI_OK;
}

// The Dirichlet mask built on the device into caller memory, with no
// allocation: mask = 1, then each included Dirichlet face multiplies its
// entries by its values, face by face in the order of build_mask, so the
// products are the same.  Each face's node list and values are copied into
// `staging` (at least nNo * (4 + 8 dof) bytes) before its kernel runs.
__global__ __launch_bounds__(256) void k_fill(size_t n, double v, double* __restrict__ x)
{
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) x[t] = v;
}

__global__ __launch_bounds__(256) void k_mask_face(int nNo, int fdof, int dof, int nd, const int* __restrict__ glob,
                                                   const double* __restrict__ val, double* __restrict__ mask)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nNo * nd) return;
    const int a = t / nd, i = t - a * nd;
    mask[(size_t)glob[a] * dof + i] *= val[(size_t)a * fdof + i];
}

int launch_fill(size_t n, double v, double* x, hipStream_t s)
{
    if (n == 0) return SVFSI_OK;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, v, x);
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

int build_mask_into(const Lhs& L, int dof, const int* incL, double* mask, void* staging, size_t stagingBytes,
                    hipStream_t s)
{
    SV_TRY(launch_fill((size_t)L.nNo * dof, 1.0, mask, s));
    for (size_t f = 0; f < L.faces.size(); ++f) {
        const Face& F = L.faces[f];
        if (!F.foC || (incL && !incL[f]) || F.bGrp != BC_TYPE_Dir || F.nNo == 0) continue;
        const int nd = F.dof < dof ? F.dof : dof;
        const size_t gb = (size_t)F.nNo * sizeof(int), vb = (size_t)F.nNo * F.dof * sizeof(double);
        const size_t voff = (gb + 255) & ~size_t(255);
        SV_ARG(voff + vb <= stagingBytes, "Dirichlet face %zu does not fit the staging buffer", f + 1);
        int* dglob = static_cast<int*>(staging);
        double* dval = reinterpret_cast<double*>(static_cast<char*>(staging) + voff);
        SV_HIP(hipMemcpyAsync(dglob, F.glob.data(), gb, hipMemcpyHostToDevice, s));
        SV_HIP(hipMemcpyAsync(dval, F.val.data(), vb, hipMemcpyHostToDevice, s));
        const int n = F.nNo * nd;
        hipLaunchKernelGGL(k_mask_face, dim3((n + 255) / 256), dim3(256), 0, s, F.nNo, F.dof, dof, nd,
                           (const int*)dglob, (const double*)dval, mask);
        SV_HIP(hipGetLastError());
    }
    return SVFSI_OK;
}

int precond_diag(Lhs& L, int dof, double* Val, double* R, double* W, float* Val32,
                 const int* incL, hipStream_t s)
{
    SV_TRY(build_mask(L, dof, incL, s));
    return precond_diag_masked(L, dof, Val, R, W, Val32, L.ws.mask.as<double>(), s);
}

int precond_diag_masked(const Lhs& L, int dof, double* Val, double* R, double* W, float* Val32,
                        const double* mask, hipStream_t s)
{
    const int n = L.nNo * dof;
    if (n == 0) return SVFSI_OK;
    hipLaunchKernelGGL(k_diag_w, dim3((n + 255) / 256), dim3(256), 0, s, L.nNo, dof,
                       L.diagPtr.as<int>(), Val, mask, W);
    const size_t nEnt = (size_t)L.nnz * dof * dof;
    const unsigned g = (unsigned)((nEnt + 255) / 256);
    switch (dof) {
    case 2: hipLaunchKernelGGL(k_scale_val<2>, dim3(g), dim3(256), 0, s, nEnt, L.blockRow.as<int>(), L.colPtr.as<int>(), W, Val, Val32); break;
    case 3: hipLaunchKernelGGL(k_scale_val<3>, dim3(g), dim3(256), 0, s, nEnt, L.blockRow.as<int>(), L.colPtr.as<int>(), W, Val, Val32); break;
    case 4: hipLaunchKernelGGL(k_scale_val<4>, dim3(g), dim3(256), 0, s, nEnt, L.blockRow.as<int>(), L.colPtr.as<int>(), W, Val, Val32); break;
    default: set_error("precond: dof %d not supported (2..4)", dof); return SVFSI_ERR_NOTSUP;
    }
    SV_TRY(launch_scale_vec((size_t)n, W, R, s));
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

// ---------------------------------------------------------------------------
// PRECONDRCS (svFSILS/PRECOND.f:150-368): Dirichlet rows and columns
// removed and their diagonal set to one, then up to 10 sweeps of row and
// column max-norm scaling; the sweep loop runs on the host (one 2-double
// read-back per sweep, as the reference's one allgather per sweep).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rcs_mask(double m) { return m > 0.5 ? 1.0 : 0.0; }

// Val <- diag(m) Val diag(m) (PREMUL then POSMUL), m the renormalised mask.
template <int DOF>
__global__ __launch_bounds__(256) void k_rcs_kill(size_t nEnt, const int* __restrict__ blockRow,
                                                  const int* __restrict__ colPtr, const double* __restrict__ mask,
                                                  double* __restrict__ val)
{
    constexpr int DD = DOF * DOF;
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nEnt) return;
    const size_t k = t / DD;
    const int e = (int)(t - k * DD);
    const int i = e / DOF, j = e - (e / DOF) * DOF;
    const double mr = rcs_mask(mask[(size_t)blockRow[k] * DOF + i]);
    const double mc = rcs_mask(mask[(size_t)colPtr[k] * DOF + j]);
    val[t] = (val[t] * mr) * mc;
}

// diagonal of masked dofs set to one, R <- m R, W1 = W2 = 1.
template <int DOF>
__global__ __launch_bounds__(256) void k_rcs_diag(int nNo, const int* __restrict__ diagPtr,
                                                  const double* __restrict__ mask, double* __restrict__ val,
                                                  double* __restrict__ R, double* __restrict__ W1,
                                                  double* __restrict__ W2)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nNo * DOF) return;
    const int a = t / DOF, i = t - (t / DOF) * DOF;
    const double m = rcs_mask(mask[t]);
    double* v = val + (size_t)diagPtr[a] * DOF * DOF + i * DOF + i;
    *v = m * (*v - 1.0) + 1.0;
    R[t] = m * R[t];
    W1[t] = 1.0;
    W2[t] = 1.0;
}

// Wr = largest |entry| of each scalar row, Wc of each scalar column (the
// blocks of column a are the transposes' of row a), and per block of
// threads max |1 - Wr|, max |1 - Wc|.
template <int DOF>
__global__ __launch_bounds__(256) void k_rcs_max(int nNo, const int* __restrict__ rowPtr,
                                                 const int* __restrict__ trans, const double* __restrict__ val,
                                                 double* __restrict__ Wr, double* __restrict__ Wc,
                                                 double* __restrict__ part)
{
    constexpr int DD = DOF * DOF;
    __shared__ double sr[256], sc[256];
    const int t = blockIdx.x * 256 + threadIdx.x;
    double dr = 0.0, dc = 0.0;
    if (t < nNo * DOF) {
        const int a = t / DOF, i = t - (t / DOF) * DOF;
        double r = 0.0, c = 0.0;
        for (int k = rowPtr[a]; k < rowPtr[a + 1]; ++k) {
            const double* b = val + (size_t)k * DD + i * DOF;
            const double* bt = val + (size_t)trans[k] * DD + i;
#pragma unroll
            for (int m = 0; m < DOF; ++m) {
                r = fmax(r, fabs(b[m]));
                c = fmax(c, fabs(bt[m * DOF]));
            }
        }
        Wr[t] = r;
        Wc[t] = c;
        dr = fabs(1.0 - r);
        dc = fabs(1.0 - c);
    }
    sr[threadIdx.x] = dr;
    sc[threadIdx.x] = dc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + o]);
            sc[threadIdx.x] = fmax(sc[threadIdx.x], sc[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sr[0];
        part[2 * blockIdx.x + 1] = sc[0];
    }
}

__global__ __launch_bounds__(256) void k_rcs_maxsum(int nb, const double* __restrict__ part, double* __restrict__ out)
{
    __shared__ double sr[256], sc[256];
    double r = 0.0, c = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        r = fmax(r, part[2 * b]);
        c = fmax(c, part[2 * b + 1]);
    }
    sr[threadIdx.x] = r;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + o]);
            sc[threadIdx.x] = fmax(sc[threadIdx.x], sc[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sr[0];
        out[1] = sc[0];
    }
}

// Wr <- 1/sqrt(Wr), Wc <- 1/sqrt(Wc), W1 *= Wr, W2 *= Wc.
__global__ __launch_bounds__(256) void k_rcs_inv(size_t n, double* __restrict__ Wr, double* __restrict__ Wc,
                                                 double* __restrict__ W1, double* __restrict__ W2)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const double r = 1.0 / sqrt(Wr[q]), c = 1.0 / sqrt(Wc[q]);
    Wr[q] = r;
    Wc[q] = c;
    W1[q] = W1[q] * r;
    W2[q] = W2[q] * c;
}

// Val <- diag(Wr) Val diag(Wc) (PREMUL then POSMUL); optional FP32 copy.
template <int DOF>
__global__ __launch_bounds__(256) void k_rcs_scale(size_t nEnt, const int* __restrict__ blockRow,
                                                   const int* __restrict__ colPtr, const double* __restrict__ Wr,
                                                   const double* __restrict__ Wc, double* __restrict__ val,
                                                   float* __restrict__ val32)
{
    constexpr int DD = DOF * DOF;
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= nEnt) return;
    const size_t k = t / DD;
    const int e = (int)(t - k * DD);
    const int i = e / DOF, j = e - (e / DOF) * DOF;
    const double v = (val[t] * Wr[(size_t)blockRow[k] * DOF + i]) * Wc[(size_t)colPtr[k] * DOF + j];
    val[t] = v;
    if (val32) val32[t] = (float)v;
}

template <int DOF>
static int precond_rcs_t(Lhs& L, double* Val, double* R, double* W1, double* W2, float* Val32, hipStream_t s)
{
    constexpr int DD = DOF * DOF;
    const size_t n = (size_t)L.nNo * DOF;
    const size_t nEnt = (size_t)L.nnz * DD;
    const unsigned ge = (unsigned)((nEnt + 255) / 256);
    const unsigned gn = (unsigned)((n + 255) / 256);
    Work& w = L.ws;
    SV_TRY(w.rcs.ensure((2 * n + 2 * (size_t)gn + 2) * sizeof(double)));
    double* Wr = w.rcs.as<double>();
    double* Wc = Wr + n;
    double* part = Wc + n;
    double* out = part + 2 * (size_t)gn;
    const double* mask = L.ws.mask.as<double>();
    hipLaunchKernelGGL(k_rcs_kill<DOF>, dim3(ge), dim3(256), 0, s, nEnt, L.blockRow.as<int>(), L.colPtr.as<int>(),
                       mask, Val);
    hipLaunchKernelGGL(k_rcs_diag<DOF>, dim3(gn), dim3(256), 0, s, L.nNo, L.diagPtr.as<int>(), mask, Val, R, W1, W2);
    for (int iter = 1;; ++iter) {
        bool last = iter >= 10;
        hipLaunchKernelGGL(k_rcs_max<DOF>, dim3(gn), dim3(256), 0, s, L.nNo, L.rowPtr.as<int>(), L.trans.as<int>(),
                           (const double*)Val, Wr, Wc, part);
        hipLaunchKernelGGL(k_rcs_maxsum, dim3(1), dim3(256), 0, s, (int)gn, (const double*)part, out);
        double h[2];
        SV_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
        SV_HIP(hipStreamSynchronize(s));
        if (h[0] < 2.0 && h[1] < 2.0) last = true;
        hipLaunchKernelGGL(k_rcs_inv, dim3(gn), dim3(256), 0, s, n, Wr, Wc, W1, W2);
        hipLaunchKernelGGL(k_rcs_scale<DOF>, dim3(ge), dim3(256), 0, s, nEnt, L.blockRow.as<int>(),
                           L.colPtr.as<int>(), (const double*)Wr, (const double*)Wc, Val, last ? Val32 : nullptr);
        if (last) break;
    }
    SV_TRY(launch_scale_vec(n, W1, R, s));
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

int precond_rcs(Lhs& L, int dof, double* Val, double* R, double* W1, double* W2, float* Val32, const int* incL,
                hipStream_t s)
{
    if (!L.symmetric) {
        set_error("PRECOND_RCS needs a structurally symmetric pattern (column maxima via the transpose map)");
        return SVFSI_ERR_NOTSUP;
    }
    SV_TRY(build_mask(L, dof, incL, s));
    if (L.nNo == 0) return SVFSI_OK;
    switch (dof) {
    case 2: return precond_rcs_t<2>(L, Val, R, W1, W2, Val32, s);
    case 3: return precond_rcs_t<3>(L, Val, R, W1, W2, Val32, s);
    case 4: return precond_rcs_t<4>(L, Val, R, W1, W2, Val32, s);
    default: set_error("PRECOND_RCS: dof %d not supported (2..4)", dof); return SVFSI_ERR_NOTSUP;
    }
}

// r = b - K x (TRANS: b - K^T x) with each row's sum carried in twice the
// working precision and rounded once: Ogita, Rump and Oishi's Dot2 ("Accurate
// sum and dot product", SIAM J. Sci. Comput. 26, 2005) -- every product
// split exactly by an FMA (a x = p + e), the running sum by Knuth's TwoSum,
// the error terms summed apart.  The computed residual is then as accurate
// as one rounded from a computation in double-double, which is what
// iterative refinement needs to reach a solution accurate to the working
// precision on an ill-conditioned system.  One thread per row component.
template <int DOF, bool TRANS>
__global__ __launch_bounds__(256) void k_resid_dd(int nNo, const int* __restrict__ rowPtr,
                                                  const int* __restrict__ colPtr, const int* __restrict__ trans,
                                                  const double* __restrict__ V, const double* __restrict__ x,
                                                  const double* __restrict__ b, double* __restrict__ r)
{
#pragma clang fp contract(off)
    constexpr int DD = DOF * DOF;
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int row = t / DOF;
    if (row >= nNo) return;
    const int p = t - row * DOF;
    double hi = b[(size_t)row * DOF + p], lo = 0.0;
    for (int k = rowPtr[row]; k < rowPtr[row + 1]; ++k) {
        const int c = colPtr[k];
        const double* blk = V + (size_t)(TRANS ? trans[k] : k) * DD;
#pragma unroll
        for (int q = 0; q < DOF; ++q) {
            const double a = TRANS ? blk[q * DOF + p] : blk[p * DOF + q];
            const double xv = x[(size_t)c * DOF + q];
            const double pr = -(a * xv);
            const double pe = __builtin_fma(-a, xv, -pr); // -a xv = pr + pe exactly
            const double sum = hi + pr;                    // TwoSum(hi, pr)
            const double bb = sum - hi;
            const double se = (hi - (sum - bb)) + (pr - bb);
            hi = sum;
            lo += se + pe;
        }
    }
    r[(size_t)row * DOF + p] = hi + lo;
}

int launch_resid_dd(const Lhs& L, int dof, const double* K, const double* x, const double* b, double* r,
                    bool transpose, hipStream_t s)
{
    if (L.nNo == 0) return SVFSI_OK;
    if (transpose && !L.symmetric) {
        set_error("transpose residual needs a structurally symmetric pattern");
        return SVFSI_ERR_ARG;
    }
    const dim3 g((unsigned)(((size_t)L.nNo * dof + 255) / 256));
    const int* rp = L.rowPtr.as<int>();
    const int* cp = L.colPtr.as<int>();
    const int* tr = transpose ? L.trans.as<int>() : nullptr;
#define SV_RDD(D)                                                                                          \
    do {                                                                                                   \
        if (transpose)                                                                                     \
            hipLaunchKernelGGL((k_resid_dd<D, true>), g, dim3(256), 0, s, L.nNo, rp, cp, tr, K, x, b, r);  \
        else                                                                                               \
            hipLaunchKernelGGL((k_resid_dd<D, false>), g, dim3(256), 0, s, L.nNo, rp, cp, tr, K, x, b, r); \
    } while (0)
    switch (dof) {
    case 2: SV_RDD(2); break;
    case 3: SV_RDD(3); break;
    case 4: SV_RDD(4); break;
    default: set_error("resid_dd: dof %d not supported (2..4)", dof); return SVFSI_ERR_NOTSUP;
    }
#undef SV_RDD
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

// Values of K^T on the (structurally symmetric) pattern of K:
// VT[k](i,j) = V[trans[k]](j,i).
template <int DOF>
__global__ __launch_bounds__(256) void k_transpose_blocks(size_t nEnt, const int* __restrict__ trans,
                                                          const double* __restrict__ V, double* __restrict__ VT)
{
    constexpr int DD = DOF * DOF;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < nEnt; q += (size_t)gridDim.x * 256) {
        const size_t blk = q / DD;
        const int e = (int)(q - blk * DD);
        const int i = e / DOF, j = e - (e / DOF) * DOF;
        VT[q] = V[(size_t)trans[blk] * DD + j * DOF + i];
    }
}

int launch_transpose_blocks(const Lhs& L, int dof, const double* V, double* VT, hipStream_t s)
{
    if (!L.symmetric) {
        set_error("transpose needs a structurally symmetric pattern");
        return SVFSI_ERR_ARG;
    }
    const size_t nEnt = (size_t)L.nnz * dof * dof;
    size_t g = (nEnt + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    const int* tr = L.trans.as<int>();
    switch (dof) {
    case 2: hipLaunchKernelGGL(k_transpose_blocks<2>, dim3((unsigned)g), dim3(256), 0, s, nEnt, tr, V, VT); break;
    case 3: hipLaunchKernelGGL(k_transpose_blocks<3>, dim3((unsigned)g), dim3(256), 0, s, nEnt, tr, V, VT); break;
    case 4: hipLaunchKernelGGL(k_transpose_blocks<4>, dim3((unsigned)g), dim3(256), 0, s, nEnt, tr, V, VT); break;
    default: set_error("transpose: dof %d not supported (2..4)", dof); return SVFSI_ERR_NOTSUP;
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

} // namespace svfsi
