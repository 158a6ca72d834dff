// amg.hip -- aggregation multigrid as the inner preconditioner of the nested
// (mixed-precision) solve; no reference counterpart (svFSILS has diagonal
// and row/column scaling only).
//
// Hierarchy: plain (unsmoothed) aggregation of the node graph, so that the
// prolongation P maps each coarse node's dof-vector to every member node
// unchanged (the dof x dof identity, the near-null space of translations)
// and the Galerkin operator P^T A P is a sum of the fine blocks: coarse
// block (I, J) = sum of the fine blocks (i, j) with i in I, j in J.
//   - Aggregation and the coarse patterns depend on the pattern only: built
//     on the host once per pattern (like LHSA) and kept with the lhs.
//   - Per solve, on the device: the coarse values (one thread per coarse
//     block entry, the fine blocks summed in a fixed order), the inverse
//     diagonal blocks, an estimate of the largest eigenvalue of D^-1 A per
//     level (power iteration), and the inverse of the coarsest operator
//     (dense Gauss-Jordan, one workgroup).
// V-cycle (FP32 vectors, FP32 matrices): Chebyshev smoothing of degree 2 on
// D^-1 A (D the node blocks, Saad's three-term recurrence, interval
// [0.3 hi, hi], hi = 1.2 times the estimate) before and after the coarse
// correction; restriction sums members, prolongation adds the coarse value.
#include <algorithm>
#include <cmath>
#include <memory>

#include "devutil.hpp"
#include "internal.hpp"
#include "stream.hpp"

namespace svfsi {

namespace {

constexpr int kDenseMax = 128; // largest coarsest operator (dofs) inverted densely
constexpr int kPowerSteps = 12;

} // namespace

struct AmgLevel {
    Lhs A;                       // pattern of a coarse level
    const Lhs* pat = nullptr;    // pattern used at this level (level 0: the caller's)
    int nNo = 0;
    std::vector<int> agg;        // host: node -> aggregate (next level's node)
    DBuf dAgg, mPtr, mIdx;       // node -> aggregate, aggregate members (CSR)
    DBuf gPtr, gIdx;             // next level's block -> this level's blocks (CSR)
    DBuf val;                    // FP32 values (coarse levels)
    const float* v = nullptr;    // values used at this level
    DBuf dinv;                   // FP32 inverse diagonal blocks
    DBuf bufs;                   // b, x, r, d, t: five vectors of nNo * dof floats
    float *b = nullptr, *x = nullptr, *r = nullptr, *d = nullptr, *t = nullptr;
    double hi = 0.0, lo = 0.0;   // Chebyshev interval for D^-1 A
};

struct AmgHier {
    std::vector<std::unique_ptr<AmgLevel>> lv;
    int dof = 0, nNo0 = 0, nnz0 = 0, cap = 0;
    int deg = 2;                 // Chebyshev degree of the smoothing
    const void* key = nullptr;   // the fine pattern's device rowPtr
    int nDense = 0;              // coarsest operator dofs when inverted densely (0: smoothed)
    DBuf dense, perm;            // coarsest inverse (nDense^2 FP64), pivot rows
    DBuf red;                    // reduction partials and the eigenvalue estimate
};

// ---------------------------------------------------------------------------
// Host: aggregation and coarse patterns.

// Greedy aggregation of the node graph (pattern only): nodes in order; an
// unaggregated node seeds an aggregate and takes its unaggregated
// neighbours, in column order, up to cap members; a node left alone joins
// the smallest neighbouring aggregate.  Returns the number of aggregates.
int amg_aggregate(int n, const int* rp, const int* cp, int cap, std::vector<int>& agg)
{
    agg.assign((size_t)n, -1);
    std::vector<int> size;
    int na = 0;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        agg[i] = na;
        int sz = 1;
        for (int k = rp[i]; k < rp[i + 1] && sz < cap; ++k) {
            const int j = cp[k];
            if (agg[j] < 0) {
                agg[j] = na;
                ++sz;
            }
        }
        size.push_back(sz);
        ++na;
    }
    for (int i = 0; i < n; ++i) {
        const int a = agg[i];
        if (size[a] != 1) continue;
        int best = -1;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int b = agg[cp[k]];
            if (b != a && (best < 0 || size[b] < size[best])) best = b;
        }
        if (best >= 0) {
            agg[i] = best;
            size[best] += 1;
            size[a] = 0;
        }
    }
    std::vector<int> id((size_t)na, -1);
    int m = 0;
    for (int a = 0; a < na; ++a)
        if (size[a] > 0) id[a] = m++;
    for (int i = 0; i < n; ++i) agg[i] = id[agg[i]];
    return m;
}

namespace {

struct CoarsePattern {
    std::vector<int> rp, cp, diag, mptr, midx, gptr, gidx;
};

// The pattern of P^T A P for plain aggregation: coarse row I holds every
// aggregate J reached from a member of I, columns sorted; gptr/gidx list,
// for each coarse block, the fine blocks summed into it in increasing order.
void coarse_pattern(int n, const std::vector<int>& rp, const std::vector<int>& cp, const std::vector<int>& agg,
                    int nc, CoarsePattern& c)
{
    c.mptr.assign((size_t)nc + 1, 0);
    for (int i = 0; i < n; ++i) c.mptr[agg[i] + 1] += 1;
    for (int a = 0; a < nc; ++a) c.mptr[a + 1] += c.mptr[a];
    c.midx.assign((size_t)n, 0);
    {
        std::vector<int> fill(c.mptr.begin(), c.mptr.end() - 1);
        for (int i = 0; i < n; ++i) c.midx[fill[agg[i]]++] = i;
    }
    const size_t nnz = (size_t)rp[n];
    std::vector<int> f2c(nnz, -1);
    std::vector<int> pos((size_t)nc, -1);
    std::vector<int> cols;
    c.rp.assign((size_t)nc + 1, 0);
    c.cp.clear();
    c.diag.assign((size_t)nc, -1);
    for (int I = 0; I < nc; ++I) {
        cols.clear();
        for (int m = c.mptr[I]; m < c.mptr[I + 1]; ++m) {
            const int i = c.midx[m];
            for (int k = rp[i]; k < rp[i + 1]; ++k) {
                const int J = agg[cp[k]];
                if (pos[J] != I) {
                    pos[J] = I;
                    cols.push_back(J);
                }
            }
        }
        std::sort(cols.begin(), cols.end());
        const int base = (int)c.cp.size();
        c.cp.insert(c.cp.end(), cols.begin(), cols.end());
        c.rp[I + 1] = (int)c.cp.size();
        for (int m = c.mptr[I]; m < c.mptr[I + 1]; ++m) {
            const int i = c.midx[m];
            for (int k = rp[i]; k < rp[i + 1]; ++k) {
                const int J = agg[cp[k]];
                f2c[k] = base + (int)(std::lower_bound(cols.begin(), cols.end(), J) - cols.begin());
            }
        }
        for (int k = c.rp[I]; k < c.rp[I + 1]; ++k)
            if (c.cp[k] == I) c.diag[I] = k;
    }
    const int nnzc = (int)c.cp.size();
    c.gptr.assign((size_t)nnzc + 1, 0);
    for (size_t k = 0; k < nnz; ++k) c.gptr[f2c[k] + 1] += 1;
    for (int k = 0; k < nnzc; ++k) c.gptr[k + 1] += c.gptr[k];
    c.gidx.assign(nnz, 0);
    std::vector<int> fill(c.gptr.begin(), c.gptr.end() - 1);
    for (size_t k = 0; k < nnz; ++k) c.gidx[fill[f2c[k]]++] = (int)k;
}

int upload(DBuf& d, const std::vector<int>& h)
{
    SV_TRY(d.ensure(std::max<size_t>(h.size(), 1) * sizeof(int)));
    if (!h.empty()) SV_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
    return SVFSI_OK;
}

int alloc_vectors(AmgLevel& l, int dof)
{
    const size_t n = (size_t)l.nNo * dof;
    SV_TRY(l.bufs.ensure(5 * std::max<size_t>(n, 1) * sizeof(float)));
    float* p = l.bufs.as<float>();
    l.b = p;
    l.x = p + n;
    l.r = p + 2 * n;
    l.d = p + 3 * n;
    l.t = p + 4 * n;
    return SVFSI_OK;
}

// The pattern-only part of the hierarchy.
int build_hierarchy(const Lhs& L, int dof, int cap, AmgHier& H)
{
    H.lv.clear();
    H.dof = dof;
    H.cap = cap;
    H.nNo0 = L.nNo;
    H.nnz0 = L.nnz;
    H.key = L.rowPtr.p;
    auto l0 = std::make_unique<AmgLevel>();
    l0->pat = &L;
    l0->nNo = L.nNo;
    SV_TRY(alloc_vectors(*l0, dof));
    H.lv.push_back(std::move(l0));
    while ((int)H.lv.size() < 16) {
        AmgLevel& f = *H.lv.back();
        const Lhs& P = *f.pat;
        if (f.nNo * dof <= kDenseMax) break;
        const int nc = amg_aggregate(f.nNo, P.hRowPtr.data(), P.hColPtr.data(), cap, f.agg);
        if (nc <= 0 || nc * 10 > f.nNo * 9) break; // coarsening stalled
        CoarsePattern c;
        coarse_pattern(f.nNo, P.hRowPtr, P.hColPtr, f.agg, nc, c);
        SV_TRY(upload(f.dAgg, f.agg));
        SV_TRY(upload(f.mPtr, c.mptr));
        SV_TRY(upload(f.mIdx, c.midx));
        SV_TRY(upload(f.gPtr, c.gptr));
        SV_TRY(upload(f.gIdx, c.gidx));
        auto l = std::make_unique<AmgLevel>();
        l->nNo = nc;
        l->A.nNo = l->A.gnNo = nc;
        l->A.nnz = (int)c.cp.size();
        SV_TRY(upload(l->A.rowPtr, c.rp));
        SV_TRY(upload(l->A.colPtr, c.cp));
        SV_TRY(upload(l->A.diagPtr, c.diag));
        l->A.hRowPtr = std::move(c.rp);
        l->A.hColPtr = std::move(c.cp);
        SV_TRY(build_row_tiles(l->A));
        l->pat = &l->A;
        SV_TRY(l->val.ensure((size_t)l->A.nnz * dof * dof * sizeof(float)));
        SV_TRY(alloc_vectors(*l, dof));
        H.lv.push_back(std::move(l));
    }
    const int nLast = H.lv.back()->nNo * dof;
    H.nDense = nLast <= kDenseMax ? nLast : 0;
    if (H.nDense) {
        SV_TRY(H.dense.ensure((size_t)H.nDense * H.nDense * sizeof(double)));
        SV_TRY(H.perm.ensure((size_t)H.nDense * sizeof(int)));
    }
    SV_TRY(H.red.ensure((256 + 8) * sizeof(double)));
    return SVFSI_OK;
}

// ---------------------------------------------------------------------------
// Device kernels.

// coarse block entry (kc, e) = sum of the fine blocks' entry e
__global__ __launch_bounds__(256) void k_amg_galerkin(size_t nEnt, int dd, const int* __restrict__ gPtr,
                                                      const int* __restrict__ gIdx, const float* __restrict__ vf,
                                                      float* __restrict__ vc)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nEnt) return;
    const size_t kc = q / dd;
    const int e = (int)(q - kc * dd);
    double s = 0.0;
    for (int m = gPtr[kc]; m < gPtr[kc + 1]; ++m) s += (double)vf[(size_t)gIdx[m] * dd + e];
    vc[q] = (float)s;
}

// y = D^-1 v for one node block (FP32 in, FP64 sums)
template <int DOF>
__device__ __forceinline__ void dinv_apply(const float* __restrict__ di, const float* v, double* y)
{
#pragma unroll
    for (int p = 0; p < DOF; ++p) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < DOF; ++q) s += (double)di[p * DOF + q] * (double)v[q];
        y[p] = s;
    }
}

// first Chebyshev step: d = c0 D^-1 r; x = d (setx) or x += d
template <int DOF>
__global__ __launch_bounds__(256) void k_amg_cheb_init(int nNo, const float* __restrict__ dinv,
                                                       const float* __restrict__ r, float* __restrict__ d,
                                                       float* __restrict__ x, float c0, int setx)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nNo) return;
    float rv[DOF];
    double z[DOF];
#pragma unroll
    for (int p = 0; p < DOF; ++p) rv[p] = r[(size_t)i * DOF + p];
    dinv_apply<DOF>(dinv + (size_t)i * DOF * DOF, rv, z);
#pragma unroll
    for (int p = 0; p < DOF; ++p) {
        const size_t q = (size_t)i * DOF + p;
        const float dv = (float)((double)c0 * z[p]);
        d[q] = dv;
        x[q] = setx ? dv : x[q] + dv;
    }
}

// later Chebyshev steps, t = A d: r -= t; d = c1 d + c2 D^-1 r; x += d
template <int DOF>
__global__ __launch_bounds__(256) void k_amg_cheb_step(int nNo, const float* __restrict__ dinv,
                                                       const float* __restrict__ t, float* __restrict__ r,
                                                       float* __restrict__ d, float* __restrict__ x, float c1,
                                                       float c2)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nNo) return;
    float rv[DOF];
#pragma unroll
    for (int p = 0; p < DOF; ++p) {
        const size_t q = (size_t)i * DOF + p;
        rv[p] = r[q] - t[q];
        r[q] = rv[p];
    }
    double z[DOF];
    dinv_apply<DOF>(dinv + (size_t)i * DOF * DOF, rv, z);
#pragma unroll
    for (int p = 0; p < DOF; ++p) {
        const size_t q = (size_t)i * DOF + p;
        const float dv = (float)((double)c1 * (double)d[q] + (double)c2 * z[p]);
        d[q] = dv;
        x[q] += dv;
    }
}

// The same steps fused with their products on the pattern's row tiles
// (stream_tile; the row sums are those of the CSR-stream SpMV, so the
// results are the separate kernels' bit for bit).
// r = b, d = c0 D^-1 b, x = d: the first step from x = 0 with the copy of b
template <int DOF>
__global__ __launch_bounds__(256) void k_amg_cheb_init_b(int nNo, const float* __restrict__ dinv,
                                                         const float* __restrict__ b, float* __restrict__ r,
                                                         float* __restrict__ d, float* __restrict__ x, float c0)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nNo) return;
    float rv[DOF];
    double z[DOF];
#pragma unroll
    for (int p = 0; p < DOF; ++p) {
        rv[p] = b[(size_t)i * DOF + p];
        r[(size_t)i * DOF + p] = rv[p];
    }
    dinv_apply<DOF>(dinv + (size_t)i * DOF * DOF, rv, z);
#pragma unroll
    for (int p = 0; p < DOF; ++p) {
        const size_t q = (size_t)i * DOF + p;
        const float dv = (float)((double)c0 * z[p]);
        d[q] = dv;
        x[q] = dv;
    }
}

// r = b - A x
template <int DOF, int T>
__global__ __launch_bounds__(256) void k_amg_resid_t(const int* __restrict__ tile, const int* __restrict__ rowPtr,
                                                     const int* __restrict__ colPtr, const float* __restrict__ val,
                                                     const float* __restrict__ x, const float* __restrict__ b,
                                                     float* __restrict__ r)
{
    stream_tile<float, DOF, T, false>(blockIdx.x, tile, rowPtr, colPtr, val, x, true,
                                      [&](int row, int i, float v, float, float, float, float) {
                                          const size_t q = (size_t)row * DOF + i;
                                          r[q] = b[q] - v;
                                      });
}

// t = A d; r -= t; dn = c1 d + c2 D^-1 r; x += dn.  dn is another buffer
// than d: other workgroups still gather d.  Every lane of a row reads the
// row's DOF residual entries before its own store (one wave, in order).
template <int DOF, int T>
__global__ __launch_bounds__(256) void k_amg_cheb_t(const int* __restrict__ tile, const int* __restrict__ rowPtr,
                                                    const int* __restrict__ colPtr, const float* __restrict__ val,
                                                    const float* __restrict__ dinv, const float* __restrict__ d,
                                                    float* __restrict__ dn, float* r, float* __restrict__ x,
                                                    float c1, float c2)
{
    stream_tile<float, DOF, T, false>(
        blockIdx.x, tile, rowPtr, colPtr, val, d, true, [&](int row, int i, float v, float s0, float s1, float s2,
                                                            float s3) {
            const size_t base = (size_t)row * DOF;
            float rv[DOF];
            rv[0] = r[base] - s0;
            rv[1] = r[base + 1] - s1;
            rv[2] = r[base + 2] - s2;
            if constexpr (DOF == 4) rv[DOF - 1] = r[base + DOF - 1] - s3;
            const float* di = dinv + (size_t)row * DOF * DOF + (size_t)i * DOF;
            double z = 0.0;
#pragma unroll
            for (int m = 0; m < DOF; ++m) z += (double)di[m] * (double)rv[m];
            const size_t q = base + i;
            const float ri = r[q] - v;
            const float dv = (float)((double)c1 * (double)d[q] + (double)c2 * z);
            dn[q] = dv;
            x[q] += dv;
            r[q] = ri;
        });
}

__global__ __launch_bounds__(256) void k_amg_sub(size_t n, const float* __restrict__ b, const float* __restrict__ t,
                                                 float* __restrict__ r)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) r[q] = b[q] - t[q];
}

// coarse right-hand side: the sum of the member nodes' residuals
template <int DOF>
__global__ __launch_bounds__(256) void k_amg_restrict(int nc, const int* __restrict__ mPtr,
                                                      const int* __restrict__ mIdx, const float* __restrict__ r,
                                                      float* __restrict__ bc)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nc * DOF) return;
    const int I = q / DOF, p = q - (q / DOF) * DOF;
    double s = 0.0;
    for (int m = mPtr[I]; m < mPtr[I + 1]; ++m) s += (double)r[(size_t)mIdx[m] * DOF + p];
    bc[q] = (float)s;
}

// x += P xc
template <int DOF>
__global__ __launch_bounds__(256) void k_amg_prolong(int nNo, const int* __restrict__ agg,
                                                     const float* __restrict__ xc, float* __restrict__ x)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nNo * DOF) return;
    const int i = q / DOF, p = q - (q / DOF) * DOF;
    x[q] += xc[(size_t)agg[i] * DOF + p];
}

// dense copy of the coarsest operator (FP64, row-major); a dof whose row and
// column are empty (Dirichlet) gets a unit diagonal
__global__ __launch_bounds__(256) void k_amg_densify(int nNo, int dof, const int* __restrict__ rp,
                                                     const int* __restrict__ cp, const float* __restrict__ v,
                                                     double* __restrict__ a)
{
    const int n = nNo * dof;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const int i = q / dof, p = q - (q / dof) * dof;
    double* row = a + (size_t)q * n;
    for (int j = 0; j < n; ++j) row[j] = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int c = cp[k];
        for (int e = 0; e < dof; ++e) row[c * dof + e] = (double)v[((size_t)k * dof + p) * dof + e];
    }
}

__global__ __launch_bounds__(256) void k_amg_unit_empty(int n, double* __restrict__ a)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    bool empty = true;
    for (int j = 0; j < n; ++j) empty = empty && a[(size_t)q * n + j] == 0.0 && a[(size_t)j * n + q] == 0.0;
    if (empty) a[(size_t)q * n + q] = 1.0;
}

// In-place Gauss-Jordan inversion with partial (row) pivoting, one
// workgroup of 1024 threads, n <= kDenseMax: each step swaps the pivot row
// in, scales it, eliminates the column from every other row (the eliminated
// column keeps the inverse's column); the row interchanges are undone as
// column interchanges in reverse order at the end.  A zero pivot is
// replaced by 1 (the inverse is then that of a regularised operator).
__global__ __launch_bounds__(1024) void k_amg_gj(int n, double* __restrict__ a, int* __restrict__ perm)
{
    __shared__ double fcol[kDenseMax];
    __shared__ double bestV[1024];
    __shared__ int bestI[1024];
    const int tid = threadIdx.x;
    for (int c = 0; c < n; ++c) {
        // pivot: the largest |a(i, c)|, i >= c (ties: the smallest i)
        double bv = -1.0;
        int bi = c;
        for (int i = c + tid; i < n; i += 1024) {
            const double v = fabs(a[(size_t)i * n + c]);
            if (v > bv) {
                bv = v;
                bi = i;
            }
        }
        bestV[tid] = bv;
        bestI[tid] = bi;
        __syncthreads();
        for (int s = 512; s > 0; s >>= 1) {
            if (tid < s) {
                const double o = bestV[tid + s];
                const int oi = bestI[tid + s];
                if (o > bestV[tid] || (o == bestV[tid] && oi < bestI[tid])) {
                    bestV[tid] = o;
                    bestI[tid] = oi;
                }
            }
            __syncthreads();
        }
        const int piv = bestI[0];
        if (tid == 0) perm[c] = piv;
        if (piv != c)
            for (int j = tid; j < n; j += 1024) {
                const double t = a[(size_t)c * n + j];
                a[(size_t)c * n + j] = a[(size_t)piv * n + j];
                a[(size_t)piv * n + j] = t;
            }
        __syncthreads();
        double pv = a[(size_t)c * n + c];
        if (pv == 0.0) pv = 1.0;
        const double ip = 1.0 / pv;
        for (int i = tid; i < n; i += 1024) fcol[i] = a[(size_t)i * n + c];
        __syncthreads();
        for (int j = tid; j < n; j += 1024) a[(size_t)c * n + j] = (j == c ? 1.0 : a[(size_t)c * n + j]) * ip;
        __syncthreads();
        for (int q = tid; q < n * n; q += 1024) {
            const int i = q / n, j = q - (q / n) * n;
            if (i == c) continue;
            const double base = (j == c) ? 0.0 : a[q];
            a[q] = base - fcol[i] * a[(size_t)c * n + j];
        }
        __syncthreads();
    }
    for (int c = n - 1; c >= 0; --c) {
        const int p = perm[c];
        if (p != c)
            for (int i = tid; i < n; i += 1024) {
                const double t = a[(size_t)i * n + c];
                a[(size_t)i * n + c] = a[(size_t)i * n + p];
                a[(size_t)i * n + p] = t;
            }
        __syncthreads();
    }
}

// x = A^-1 b with the dense coarsest inverse
__global__ __launch_bounds__(256) void k_amg_dense_mv(int n, const double* __restrict__ ainv,
                                                      const float* __restrict__ b, float* __restrict__ x)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += ainv[(size_t)i * n + j] * (double)b[j];
    x[i] = (float)s;
}

// power iteration pieces: a fixed pseudo-random start, the squared-norm
// partials, and v = t / |t| (every block sums the partials itself; block 0
// stores |t|)
__global__ __launch_bounds__(256) void k_amg_rand(size_t n, float* __restrict__ v)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    unsigned x = (unsigned)q * 2654435761u ^ 0x9e3779b9u;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    v[q] = (float)(x >> 8) * (1.0f / 16777216.0f) - 0.5f;
}

__global__ __launch_bounds__(256) void k_amg_sumsq(size_t n, const float* __restrict__ v, double* __restrict__ part)
{
    __shared__ double lds[4];
    double s = 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        const double a = (double)v[q];
        s += a * a;
    }
    s = block_sum<256>(s, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_amg_normalize(size_t n, const float* __restrict__ t, float* __restrict__ v,
                                                       const double* __restrict__ part, int nb,
                                                       double* __restrict__ normOut)
{
    __shared__ double tot;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += part[b];
        tot = sqrt(s);
    }
    __syncthreads();
    const double nrm = tot;
    if (blockIdx.x == 0 && threadIdx.x == 0) *normOut = nrm;
    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        v[q] = (float)((double)t[q] * inv);
}

unsigned grid_of(size_t n)
{
    return (unsigned)std::max<size_t>(1, (n + 255) / 256);
}

template <int DOF>
int apply_dinv(const AmgLevel& l, float* v, hipStream_t s)
{
    return launch_blockdiag_apply<float>(*l.pat, DOF, l.dinv.as<float>(), v, s, nullptr);
}

// profiling slot of a level's products: 1 (the inner FP32 SpMV) on the
// finest level, 5 on the coarse ones
inline int spmv_slot(const AmgLevel& l) { return l.pat == &l.A ? 5 : 1; }

// Chebyshev smoothing of A x = b on D^-1 A, degree deg: x given (setx:
// starts from zero) and r = b - A x on entry when !setx; on exit r holds
// b - A x minus the last step's update, x the smoothed iterate.
// The level's row tiles for the fused kernels (those of its CSR-stream
// SpMV), or nullptr when the pattern has none (SVFSI_AMG_FUSED=0 also
// selects the separate kernels).
inline const int* level_tiles(const AmgLevel& l, int dof, int& T, int& nT)
{
    const char* e = getenv("SVFSI_AMG_FUSED");
    const bool fused = !(e && e[0] == '0');
    return fused ? stream_tiles(*l.pat, dof, false, T, nT) : nullptr;
}

// bytes of the level's FP32 SpMV (values, indices, row pointers, x and y once)
inline double level_spmv_bytes(const AmgLevel& l, int dof)
{
    const Lhs& P = *l.pat;
    return (double)P.nnz * dof * dof * 4.0 + 4.0 * P.nnz + 4.0 * (P.nNo + 1) + 2.0 * P.nNo * dof * 4.0;
}

// r = b - A x (the residual and its product in one launch when tiled)
template <int DOF>
int level_resid(AmgLevel& l, hipStream_t s)
{
    const size_t n = (size_t)l.nNo * DOF;
    int T = 0, nT = 0;
    if (const int* tl = level_tiles(l, DOF, T, nT)) {
        ProfScope ps(spmv_slot(l), level_spmv_bytes(l, DOF) + n * 4.0);
        const int* rp = l.pat->rowPtr.as<int>();
        const int* cp = l.pat->colPtr.as<int>();
        if (T == 512)
            SV_LAUNCH(ps, (k_amg_resid_t<DOF, 512>), dim3((unsigned)nT), dim3(256), 0, s, tl, rp, cp, l.v,
                      (const float*)l.x, (const float*)l.b, l.r);
        else
            SV_LAUNCH(ps, (k_amg_resid_t<DOF, 256>), dim3((unsigned)nT), dim3(256), 0, s, tl, rp, cp, l.v,
                      (const float*)l.x, (const float*)l.b, l.r);
        SV_HIP(hipGetLastError());
        return SVFSI_OK;
    }
    SV_TRY(launch_spmv_f32(*l.pat, DOF, l.v, l.x, l.t, s, nullptr, nullptr, spmv_slot(l)));
    hipLaunchKernelGGL(k_amg_sub, dim3(grid_of(n)), dim3(256), 0, s, n, (const float*)l.b, (const float*)l.t, l.r);
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

template <int DOF>
int cheb(AmgLevel& l, int deg, bool setx, hipStream_t s, bool fromB = false)
{
    const double th = 0.5 * (l.hi + l.lo), de = 0.5 * (l.hi - l.lo);
    const double s1 = th / de;
    double rho0 = 1.0 / s1;
    const float* dinv = l.dinv.as<float>();
    const size_t n = (size_t)l.nNo * DOF;
    if (fromB)
        hipLaunchKernelGGL(k_amg_cheb_init_b<DOF>, dim3(grid_of(l.nNo)), dim3(256), 0, s, l.nNo, dinv,
                           (const float*)l.b, l.r, l.d, l.x, (float)(1.0 / th));
    else
        hipLaunchKernelGGL(k_amg_cheb_init<DOF>, dim3(grid_of(l.nNo)), dim3(256), 0, s, l.nNo, dinv,
                           (const float*)l.r, l.d, l.x, (float)(1.0 / th), setx ? 1 : 0);
    int T = 0, nT = 0;
    const int* tl = level_tiles(l, DOF, T, nT);
    for (int k = 1; k < deg; ++k) {
        const double rho1 = 1.0 / (2.0 * s1 - rho0);
        const float c1 = (float)(rho1 * rho0), c2 = (float)(2.0 * rho1 / de);
        if (tl) {
            ProfScope ps(spmv_slot(l), level_spmv_bytes(l, DOF) + n * 4.0 * 5.0 + (double)l.nNo * DOF * DOF * 4.0);
            const int* rp = l.pat->rowPtr.as<int>();
            const int* cp = l.pat->colPtr.as<int>();
            if (T == 512)
                SV_LAUNCH(ps, (k_amg_cheb_t<DOF, 512>), dim3((unsigned)nT), dim3(256), 0, s, tl, rp, cp, l.v, dinv,
                          (const float*)l.d, l.t, l.r, l.x, c1, c2);
            else
                SV_LAUNCH(ps, (k_amg_cheb_t<DOF, 256>), dim3((unsigned)nT), dim3(256), 0, s, tl, rp, cp, l.v, dinv,
                          (const float*)l.d, l.t, l.r, l.x, c1, c2);
            std::swap(l.d, l.t); // the new direction is in t
        } else {
            SV_TRY(launch_spmv_f32(*l.pat, DOF, l.v, l.d, l.t, s, nullptr, nullptr, spmv_slot(l)));
            hipLaunchKernelGGL(k_amg_cheb_step<DOF>, dim3(grid_of(l.nNo)), dim3(256), 0, s, l.nNo, dinv,
                               (const float*)l.t, l.r, l.d, l.x, c1, c2);
        }
        rho0 = rho1;
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

template <int DOF>
int vcycle(AmgHier& H, size_t li, hipStream_t s)
{
    AmgLevel& l = *H.lv[li];
    const size_t n = (size_t)l.nNo * DOF;
    const bool last = li + 1 == H.lv.size();
    if (last) {
        if (H.nDense) {
            hipLaunchKernelGGL(k_amg_dense_mv, dim3(grid_of(n)), dim3(256), 0, s, (int)n,
                               (const double*)H.dense.as<double>(), (const float*)l.b, l.x);
        } else {
            SV_HIP(hipMemcpyAsync(l.r, l.b, n * sizeof(float), hipMemcpyDeviceToDevice, s));
            SV_TRY(cheb<DOF>(l, 16, true, s));
        }
        SV_HIP(hipGetLastError());
        return SVFSI_OK;
    }
    AmgLevel& c = *H.lv[li + 1];
    // pre-smoothing from x = 0 (r = b, copied by the first step), which
    // leaves r = b - A x current but for the last step's update
    SV_TRY(cheb<DOF>(l, H.deg, true, s, true));
    SV_TRY(level_resid<DOF>(l, s));
    hipLaunchKernelGGL(k_amg_restrict<DOF>, dim3(grid_of((size_t)c.nNo * DOF)), dim3(256), 0, s, c.nNo,
                       l.mPtr.as<int>(), l.mIdx.as<int>(), (const float*)l.r, c.b);
    SV_TRY(vcycle<DOF>(H, li + 1, s));
    hipLaunchKernelGGL(k_amg_prolong<DOF>, dim3(grid_of(n)), dim3(256), 0, s, l.nNo, l.dAgg.as<int>(),
                       (const float*)c.x, l.x);
    // post-smoothing from the corrected x
    SV_TRY(level_resid<DOF>(l, s));
    SV_TRY(cheb<DOF>(l, H.deg, false, s));
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

// largest eigenvalue of D^-1 A by power iteration (the estimate's norm read
// back once per level)
template <int DOF>
int estimate_lambda(AmgHier& H, AmgLevel& l, hipStream_t s, double& lam)
{
    const size_t n = (size_t)l.nNo * DOF;
    const unsigned nb = std::min<unsigned>(256, grid_of(n));
    double* part = H.red.as<double>();
    double* nrm = part + 256;
    hipLaunchKernelGGL(k_amg_rand, dim3(grid_of(n)), dim3(256), 0, s, n, l.d);
    for (int it = 0; it < kPowerSteps; ++it) {
        SV_TRY(launch_spmv_f32(*l.pat, DOF, l.v, l.d, l.t, s, nullptr, nullptr, spmv_slot(l)));
        SV_TRY(apply_dinv<DOF>(l, l.t, s));
        hipLaunchKernelGGL(k_amg_sumsq, dim3(nb), dim3(256), 0, s, n, (const float*)l.t, part);
        hipLaunchKernelGGL(k_amg_normalize, dim3(nb), dim3(256), 0, s, n, (const float*)l.t, l.d,
                           (const double*)part, (int)nb, nrm);
    }
    SV_HIP(hipGetLastError());
    SV_HIP(hipMemcpyAsync(&lam, nrm, sizeof(double), hipMemcpyDeviceToHost, s));
    SV_HIP(hipStreamSynchronize(s));
    return SVFSI_OK;
}

template <int DOF>
int setup_values(AmgHier& H, const double* K64, const float* K32, hipStream_t s)
{
    constexpr int DD = DOF * DOF;
    for (size_t li = 0; li < H.lv.size(); ++li) {
        AmgLevel& l = *H.lv[li];
        if (li == 0) {
            l.v = K32;
            SV_TRY(l.dinv.ensure((size_t)l.nNo * DD * sizeof(float)));
            SV_TRY(launch_blockdiag_inv<float>(*l.pat, DOF, K64, l.dinv.as<float>(), s));
        } else {
            AmgLevel& f = *H.lv[li - 1];
            const size_t nEnt = (size_t)l.A.nnz * DD;
            hipLaunchKernelGGL(k_amg_galerkin, dim3(grid_of(nEnt)), dim3(256), 0, s, nEnt, DD, f.gPtr.as<int>(),
                               f.gIdx.as<int>(), f.v, l.val.as<float>());
            l.v = l.val.as<float>();
            SV_TRY(l.dinv.ensure((size_t)l.nNo * DD * sizeof(float)));
            SV_TRY(launch_blockdiag_inv_f32(*l.pat, DOF, l.v, l.dinv.as<float>(), s));
        }
        const bool last = li + 1 == H.lv.size();
        if (!(last && H.nDense)) {
            double lam = 0.0;
            SV_TRY(estimate_lambda<DOF>(H, l, s, lam));
            if (!(lam > 0.0) || !std::isfinite(lam)) lam = 2.0;
            l.hi = 1.2 * lam;
            l.lo = (last ? 0.05 : 0.3) * l.hi;
        }
    }
    if (H.nDense) {
        AmgLevel& l = *H.lv.back();
        const int n = H.nDense;
        hipLaunchKernelGGL(k_amg_densify, dim3(grid_of(n)), dim3(256), 0, s, l.nNo, DOF, l.pat->rowPtr.as<int>(),
                           l.pat->colPtr.as<int>(), l.v, H.dense.as<double>());
        hipLaunchKernelGGL(k_amg_unit_empty, dim3(grid_of(n)), dim3(256), 0, s, n, H.dense.as<double>());
        hipLaunchKernelGGL(k_amg_gj, dim3(1), dim3(1024), 0, s, n, H.dense.as<double>(), H.perm.as<int>());
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

} // namespace

// Set-up for one solve: the hierarchy (rebuilt when the pattern or dof
// changes) and its values from the scaled matrix (K64 FP64, for the finest
// diagonal blocks; K32 its FP32 copy, the finest operator).
int amg_setup(Lhs& L, int dof, const double* K64, const float* K32, hipStream_t s)
{
    if (dof != 3 && dof != 4) {
        set_error("aggregation multigrid: dof %d not supported (3, 4)", dof);
        return SVFSI_ERR_NOTSUP;
    }
    // aggregate size cap (SVFSI_AMG_CAP): measured on the 3.2M-tet cylinder,
    // ms per mixed step at caps 4 / 5 / 6 / 7 / 8 / 12: 194 / 177 / 191 /
    // 204 / 215 / 330
    const char* e = getenv("SVFSI_AMG_CAP");
    int cap = e ? atoi(e) : 5;
    if (cap < 2) cap = 2;
    Work& w = L.ws;
    if (!w.amg || w.amg->key != L.rowPtr.p || w.amg->dof != dof || w.amg->nNo0 != L.nNo || w.amg->nnz0 != L.nnz ||
        w.amg->cap != cap) {
        if (L.hRowPtr.size() != (size_t)L.nNo + 1 || L.hColPtr.size() != (size_t)L.nnz) {
            set_error("aggregation multigrid: the pattern's host copy is missing");
            return SVFSI_ERR_ARG;
        }
        w.amg = std::shared_ptr<AmgHier>(new AmgHier());
        SV_TRY(build_hierarchy(L, dof, cap, *w.amg));
    }
    AmgHier& H = *w.amg;
    H.lv[0]->pat = &L;
    // Chebyshev degree of the pre- and post-smoothing (SVFSI_AMG_DEG)
    const char* ed = getenv("SVFSI_AMG_DEG");
    H.deg = ed ? std::max(1, std::min(8, atoi(ed))) : 2;
    return dof == 3 ? setup_values<3>(H, K64, K32, s) : setup_values<4>(H, K64, K32, s);
}

// z = M^-1 q: one V-cycle (FP32); q and z are n = nNo * dof vectors.
int amg_apply(Lhs& L, int dof, const float* q, float* z, hipStream_t s)
{
    if (!L.ws.amg) {
        set_error("aggregation multigrid: no set-up");
        return SVFSI_ERR_ARG;
    }
    AmgHier& H = *L.ws.amg;
    AmgLevel& l0 = *H.lv[0];
    const size_t n = (size_t)l0.nNo * dof;
    SV_HIP(hipMemcpyAsync(l0.b, q, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    SV_TRY(dof == 3 ? vcycle<3>(H, 0, s) : vcycle<4>(H, 0, s));
    SV_HIP(hipMemcpyAsync(z, l0.x, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    return SVFSI_OK;
}

// levels and sizes of the current hierarchy (for reports and tests)
int amg_levels(const Lhs& L, std::vector<int>& nodes)
{
    nodes.clear();
    if (!L.ws.amg) return 0;
    for (auto& l : L.ws.amg->lv) nodes.push_back(l->nNo);
    return (int)nodes.size();
}

} // namespace svfsi
