// nssolver.hip -- NSSOLVER (svFSILS/NSSOLVER.f:52-383): the reference's
// block solver for the incompressible Navier-Stokes system (dof 4 =
// velocity 3 + pressure), FP64, one device.
//
// Each outer iteration adds two directions, [0; P_i] with
// P_i = S^-1 (Rc - D K^-1 Rm) (CGRAD_SCHUR, S = L + G^T G) and [U_i; 0] with
// U_i = K^-1 (Rm - G P_i) (GMRES on the momentum block), forms their images
// under the full matrix, and solves the Gram system of all images so far
// for the least-squares combination (GE).  The block products, the vector
// updates, the dots and the Schur-complement CG run on the device; the
// outer loop and GE (at most 2 mItr unknowns) on the host, with all Gram
// entries of an outer iteration brought back in one read-back (the
// reference makes one allreduce per dot).
#include "internal.hpp"
#include "devutil.hpp"

#include <cmath>
#include <vector>

namespace svfsi {

namespace {

constexpr int kNsd = 3;

// DEPART (NSSOLVER.f:236-303): mK (9), mG (3), mD (3), mL (1) of each 4x4
// block, and Gt(:, trans(k)) = -mG(:, k).
__global__ __launch_bounds__(256) void k_depart(size_t nnz, const int* __restrict__ trans,
                                                const double* __restrict__ val, double* __restrict__ mK,
                                                double* __restrict__ mG, double* __restrict__ mD,
                                                double* __restrict__ mL, double* __restrict__ Gt)
{
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= nnz) return;
    const double* t = val + k * 16;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) mK[k * 9 + r * 3 + c] = t[r * 4 + c];
    const size_t l = (size_t)trans[k];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        mG[k * 3 + m] = t[m * 4 + 3];
        mD[k * 3 + m] = t[12 + m];
        Gt[l * 3 + m] = -t[m * 4 + 3];
    }
    mL[k] = t[15];
}

// FSILS_SPARMULVS / SV / SS (SPARMUL.f:135-297): 16 lanes per row (a tet4
// node has about 15 neighbours), lane l taking the row's blocks l, l+16, ...
// so that the loads of K and colPtr are contiguous; shuffles add the lanes.
constexpr int kLanes = 16;

__device__ __forceinline__ double lanes_sum(double v)
{
    for (int o = kLanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kLanes);
    return v;
}

__global__ __launch_bounds__(256) void k_sp_vs(int nNo, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colPtr, const double* __restrict__ K,
                                               const double* __restrict__ U, double* __restrict__ KU)
{
    const int i = blockIdx.x * (256 / kLanes) + threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    if (i >= nNo) return;
    double a = 0.0;
    for (int j = rowPtr[i] + lane; j < rowPtr[i + 1]; j += kLanes) {
        const double* k = K + (size_t)j * kNsd;
        const double* u = U + (size_t)colPtr[j] * kNsd;
        a += k[0] * u[0] + k[1] * u[1] + k[2] * u[2];
    }
    a = lanes_sum(a);
    if (lane == 0) KU[i] = a;
}

__global__ __launch_bounds__(256) void k_sp_sv(int nNo, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colPtr, const double* __restrict__ K,
                                               const double* __restrict__ U, double* __restrict__ KU)
{
    const int i = blockIdx.x * (256 / kLanes) + threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    if (i >= nNo) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int j = rowPtr[i] + lane; j < rowPtr[i + 1]; j += kLanes) {
        const double u = U[colPtr[j]];
        a0 += K[(size_t)j * 3] * u;
        a1 += K[(size_t)j * 3 + 1] * u;
        a2 += K[(size_t)j * 3 + 2] * u;
    }
    a0 = lanes_sum(a0);
    a1 = lanes_sum(a1);
    a2 = lanes_sum(a2);
    if (lane == 0) {
        KU[(size_t)i * 3] = a0;
        KU[(size_t)i * 3 + 1] = a1;
        KU[(size_t)i * 3 + 2] = a2;
    }
}

__global__ __launch_bounds__(256) void k_sp_ss(int nNo, const int* __restrict__ rowPtr,
                                               const int* __restrict__ colPtr, const double* __restrict__ K,
                                               const double* __restrict__ U, double* __restrict__ KU)
{
    const int i = blockIdx.x * (256 / kLanes) + threadIdx.x / kLanes;
    const int lane = threadIdx.x % kLanes;
    if (i >= nNo) return;
    double a = 0.0;
    for (int j = rowPtr[i] + lane; j < rowPtr[i + 1]; j += kLanes) a += K[j] * U[colPtr[j]];
    a = lanes_sum(a);
    if (lane == 0) KU[i] = a;
}

// y <- a*y + c*x (a = 1: the reference's OMPSUMS/V, y = y + c x; a = 0:
// y = c x), and z <- x - y.
__global__ __launch_bounds__(256) void k_ns_axpy(size_t n, double a, double* __restrict__ y, double c,
                                                 const double* __restrict__ x)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    y[q] = (a == 0.0) ? c * x[q] : y[q] + c * x[q];
}

__global__ __launch_bounds__(256) void k_ns_sub(size_t n, const double* __restrict__ x,
                                                const double* __restrict__ y, double* __restrict__ z)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) z[q] = x[q] - y[q];
}

// Ri (4, nNo) <-> Rm (3, nNo), Rc (nNo)
__global__ __launch_bounds__(256) void k_ns_split(int nNo, const double* __restrict__ Ri, double* __restrict__ Rm,
                                                  double* __restrict__ Rc)
{
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= nNo) return;
    for (int m = 0; m < 3; ++m) Rm[(size_t)a * 3 + m] = Ri[(size_t)a * 4 + m];
    Rc[a] = Ri[(size_t)a * 4 + 3];
}

__global__ __launch_bounds__(256) void k_ns_merge(int nNo, const double* __restrict__ Rm,
                                                  const double* __restrict__ Rc, double* __restrict__ Ri)
{
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= nNo) return;
    for (int m = 0; m < 3; ++m) Ri[(size_t)a * 4 + m] = Rm[(size_t)a * 3 + m];
    Ri[(size_t)a * 4 + 3] = Rc[a];
}

inline dim3 grid_of(size_t n) { return dim3((unsigned)((n + 255) / 256)); }
inline dim3 row_grid(size_t nRows) { return dim3((unsigned)((nRows + 256 / kLanes - 1) / (256 / kLanes))); }

// The Gram entries of one outer iteration (NSSOLVER.f:135-146) in one pass:
// pair p is (k, j) with k = kLo for p <= kLo, else k = kLo + 1, and j = 0..k
// (j = 0: the right-hand side Rmi, Rci).  Each pair gets block partials of
// MU_j . MU_k over 3 nNo values (slot 2p) and of MP_j . MP_k over nNo
// values (slot 2p+1); k_ns_gram_sum adds each slot's partials in a fixed
// order.
constexpr int kGramBlocks = 64;

__global__ __launch_bounds__(256) void k_ns_gram(int nNo, int kLo, const double* __restrict__ MU,
                                                 const double* __restrict__ MP, const double* __restrict__ Rmi,
                                                 const double* __restrict__ Rci, double* __restrict__ part)
{
    __shared__ double lds[4];
    const int p = blockIdx.y;
    const int k = p <= kLo ? kLo : kLo + 1;
    const int j = p <= kLo ? p : p - kLo - 1;
    const size_t n3 = (size_t)nNo * kNsd;
    const double* a = j == 0 ? Rmi : MU + (size_t)(j - 1) * n3;
    const double* b = MU + (size_t)(k - 1) * n3;
    const double* ap = j == 0 ? Rci : MP + (size_t)(j - 1) * nNo;
    const double* bp = MP + (size_t)(k - 1) * nNo;
    double su = 0.0, sp = 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n3; q += (size_t)gridDim.x * 256) su += a[q] * b[q];
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < (size_t)nNo; q += (size_t)gridDim.x * 256)
        sp += ap[q] * bp[q];
    su = block_sum<256>(su, lds);
    sp = block_sum<256>(sp, lds);
    if (threadIdx.x == 0) {
        part[(size_t)(2 * p) * gridDim.x + blockIdx.x] = su;
        part[(size_t)(2 * p + 1) * gridDim.x + blockIdx.x] = sp;
    }
}

__global__ __launch_bounds__(64) void k_ns_gram_sum(int gx, const double* __restrict__ part, double* __restrict__ out)
{
    double v = 0.0;
    for (int q = threadIdx.x; q < gx; q += 64) v += part[(size_t)blockIdx.x * gx + q];
    v = wave_sum(v);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// GE (svFSILS/GE.f:51-151), host: A column-major with leading dimension nV;
// B overwritten with the solution; false (and B = 0) if singular.
bool ge_host(int nV, int N, const std::vector<double>& A, std::vector<double>& B)
{
    auto AA = [&](int i, int j) { return A[(size_t)j * nV + i]; };
    if (N <= 0) return false;
    std::vector<double> W(N);
    for (int i = 0; i < N; ++i) {
        if (std::fabs(AA(i, i)) < 2.2250738585072014e-308) {
            for (int q = 0; q < N; ++q) B[q] = 0.0;
            return false;
        }
        W[i] = 1.0 / std::sqrt(std::fabs(AA(i, i)));
    }
    std::vector<double> C((size_t)N * (N + 1));
    auto CC = [&](int i, int j) -> double& { return C[(size_t)i * (N + 1) + j]; };
    for (int i = 0; i < N; ++i) {
        for (int j = 0; j < N; ++j) CC(i, j) = W[i] * W[j] * AA(i, j);
        CC(i, N) = W[i] * B[i];
    }
    const double eps = 2.220446049250313e-16;
    if (N == 1) {
        B[0] = CC(0, 1) / CC(0, 0);
        B[0] = B[0] * W[0];
        return true;
    }
    if (N == 2) {
        const double pivot = CC(0, 0) * CC(1, 1) - CC(1, 0) * CC(0, 1);
        if (std::fabs(pivot) < eps) {
            B[0] = B[1] = 0.0;
            return false;
        }
        B[0] = (CC(0, 2) * CC(1, 1) - CC(1, 2) * CC(0, 1)) / pivot;
        B[1] = (CC(1, 2) * CC(0, 0) - CC(0, 2) * CC(1, 0)) / pivot;
        B[0] = W[0] * B[0];
        B[1] = W[1] * B[1];
        return true;
    }
    for (int m = 0; m < N - 1; ++m) {
        int ipv = m;
        double pivot = std::fabs(CC(m, m));
        for (int i = m + 1; i < N; ++i)
            if (std::fabs(CC(i, m)) > pivot) {
                ipv = i;
                pivot = std::fabs(CC(i, m));
            }
        if (std::fabs(pivot) < eps) {
            for (int q = 0; q < N; ++q) B[q] = 0.0;
            return false;
        }
        if (ipv != m)
            for (int j = m; j <= N; ++j) std::swap(CC(m, j), CC(ipv, j));
        for (int i = m + 1; i < N; ++i) {
            const double f = CC(i, m) / CC(m, m);
            CC(i, m) = 0.0;
            for (int j = m + 1; j <= N; ++j) CC(i, j) = CC(i, j) - f * CC(m, j);
        }
    }
    for (int j = N - 1; j >= 0; --j) {
        for (int i = j + 1; i < N; ++i) CC(j, N) = CC(j, N) - CC(j, i) * CC(i, N);
        CC(j, N) = CC(j, N) / CC(j, j);
    }
    for (int i = 0; i < N; ++i) B[i] = W[i] * CC(i, N);
    return true;
}

struct NsCtx {
    Lhs* L;
    hipStream_t s;
    size_t n;       // nodes
    double* part;   // dot partials + result
    double dot(const double* u, const double* v, size_t len, int* rc)
    {
        double out = 0.0;
        int r = launch_dot((int)len, u, v, part, part + 512, s);
        if (r == SVFSI_OK && hipMemcpyAsync(&out, part + 512, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess)
            r = SVFSI_ERR_HIP;
        if (r == SVFSI_OK && hipStreamSynchronize(s) != hipSuccess) r = SVFSI_ERR_HIP;
        if (r != SVFSI_OK) *rc = r;
        return out;
    }
    void axpy(double* y, double a, double c, const double* x, size_t len)
    {
        hipLaunchKernelGGL(k_ns_axpy, grid_of(len), dim3(256), 0, s, len, a, y, c, x);
    }
    void vs(const double* K, const double* U, double* KU)
    {
        hipLaunchKernelGGL(k_sp_vs, row_grid(n), dim3(256), 0, s, (int)n, L->rowPtr.as<int>(), L->colPtr.as<int>(),
                           K, U, KU);
    }
    void sv(const double* K, const double* U, double* KU)
    {
        hipLaunchKernelGGL(k_sp_sv, row_grid(n), dim3(256), 0, s, (int)n, L->rowPtr.as<int>(), L->colPtr.as<int>(),
                           K, U, KU);
    }
    void ss(const double* K, const double* U, double* KU)
    {
        hipLaunchKernelGGL(k_sp_ss, row_grid(n), dim3(256), 0, s, (int)n, L->rowPtr.as<int>(), L->colPtr.as<int>(),
                           K, U, KU);
    }
};

// GMRES (svFSILS/GMRES.f:51-166) for K X = R (dof 3) through the FP64
// GMRESV driver on a copy of R.  The two differ only in bookkeeping: GMRES
// takes its first cycle from R without the product K 0 (the same numbers)
// and counts one iteration fewer per call; its early return (|R| <= absTol)
// leaves X = 0 where GMRESV leaves R.  itr and callD accumulate.
int ns_gmres(NsCtx& c, const double* K, const double* R, double* X, FSILS_subLsType& gm)
{
    const size_t len = c.n * kNsd;
    SV_HIP(hipMemcpyAsync(X, R, len * sizeof(double), hipMemcpyDeviceToDevice, c.s));
    // coupled faces: GMRES on PRE (K + ADD) X = PRE R, PRE being linear
    // (GMRES.f:79-115 applies PRE to each residual and product)
    const bool cpl = c.L->cpl.active;
    if (cpl) SV_TRY(cpl_apply(*c.L, 1, kNsd, X, X, c.s));
    FSILS_subLsType t = gm;
    t.itr = 0;
    SV_TRY(gmres_solve(*c.L, t, kNsd, K, X, c.s, false, cpl ? 2 : 0));
    if (t.iNorm <= gm.absTol) {
        SV_HIP(hipMemsetAsync(X, 0, len * sizeof(double), c.s));
        gm.callD = 2.220446049250313e-16;  // GMRES.f:97: assigned, as the reference does
        gm.dB = 0.0;
        return SVFSI_OK;
    }
    const int add = t.itr > 0 ? t.itr - 1 : 0;
    const double callD = gm.callD + t.callD;
    const int itr = gm.itr + add;
    gm = t;
    gm.itr = itr;
    gm.callD = callD;
    return SVFSI_OK;
}

} // namespace

int ns_solve(Lhs& L, FSILS_lsType& ls, int dof, const double* Val, double* Ri, hipStream_t s)
{
    if (dof != 4) {
        set_error("NS solver: dof %d not supported (4: velocity 3 + pressure)", dof);
        return SVFSI_ERR_NOTSUP;
    }
    if (!L.symmetric) {
        set_error("NS solver needs a structurally symmetric pattern (the G^T blocks)");
        return SVFSI_ERR_NOTSUP;
    }
    const size_t n = (size_t)L.nNo, nnz = (size_t)L.nnz;
    const int iBmax = ls.RI.mItr, nB = 2 * iBmax;
    if (iBmax < 1) {
        set_error("NS solver: RI.mItr must be positive");
        return SVFSI_ERR_ARG;
    }
    Work& w = L.ws;
    const size_t need = 19 * nnz + n * (3 + 1 + 3 + 1) + n * (3 + 1) * (size_t)iBmax + n * (3 + 1) * (size_t)nB +
                        n * (1 + 1 + 1 + 1 + 1 + 3) + 520 + (size_t)2 * (2 * nB + 1) * (kGramBlocks + 1);
    SV_TRY(w.ns.ensure(need * sizeof(double)));
    double* p = w.ns.as<double>();
    auto take = [&](size_t k) { double* q = p; p += k; return q; };
    double* mK = take(9 * nnz);
    double* mG = take(3 * nnz);
    double* mD = take(3 * nnz);
    double* mL = take(nnz);
    double* Gt = take(3 * nnz);
    double* Rm = take(3 * n);
    double* Rc = take(n);
    double* Rmi = take(3 * n);
    double* Rci = take(n);
    double* U = take(3 * n * iBmax);
    double* P = take(n * iBmax);
    double* MU = take(3 * n * nB);
    double* MP = take(n * nB);
    // the Schur CG's right-hand side / solution at one fixed address, so its
    // captured iteration graphs are reused across outer iterations
    double* cR = take(n);
    double* cX = take(n);
    double* cP = take(n);
    double* cP2 = take(n);
    double* cSP = take(n);
    double* cGP = take(3 * n);
    NsCtx c{&L, s, n, take(520)};
    const int npMax = 2 * nB + 1;  // pairs of the last outer iteration
    double* gpart = take((size_t)2 * npMax * kGramBlocks);
    double* gout = take((size_t)2 * npMax);
    std::vector<double> gram(2 * npMax);
    auto VU = [&](int i) { return U + (size_t)(i - 1) * 3 * n; };
    auto VP = [&](int i) { return P + (size_t)(i - 1) * n; };
    auto VMU = [&](int i) { return MU + (size_t)(i - 1) * 3 * n; };
    auto VMP = [&](int i) { return MP + (size_t)(i - 1) * n; };
    int rc = SVFSI_OK;

    hipLaunchKernelGGL(k_ns_split, grid_of(n), dim3(256), 0, s, (int)n, (const double*)Ri, Rmi, Rci);
    SV_HIP(hipMemcpyAsync(Rm, Rmi, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    SV_HIP(hipMemcpyAsync(Rc, Rci, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    const double nm = std::sqrt(c.dot(Rm, Rm, 3 * n, &rc));
    const double nc = std::sqrt(c.dot(Rc, Rc, n, &rc));
    SV_TRY(rc);
    double eps = std::sqrt(nm * nm + nc * nc);
    ls.RI.iNorm = eps;
    ls.RI.fNorm = eps * eps;
    ls.CG.callD = 0.0;
    ls.GM.callD = 0.0;
    ls.CG.itr = 0;
    ls.GM.itr = 0;
    ls.RI.callD = wall_seconds();
    ls.RI.suc = 0;
    eps = std::fmax(ls.RI.absTol, ls.RI.relTol * eps);
    hipLaunchKernelGGL(k_depart, grid_of(nnz), dim3(256), 0, s, nnz, L.trans.as<int>(), Val, mK, mG, mD, mL, Gt);
    SV_HIP(hipGetLastError());

    std::vector<double> A((size_t)nB * nB, 0.0), B(nB, 0.0), xB(nB, 0.0), oldxB(nB, 0.0);
    auto AM = [&](int j, int k) -> double& { return A[(size_t)(k - 1) * nB + (j - 1)]; };
    int i, iB = 0, iBB = 0;
    for (i = 1; i <= ls.RI.mItr; ++i) {
        iB = 2 * i - 1;
        iBB = 2 * i;
        ls.RI.dB = ls.RI.fNorm;
        SV_TRY(ns_gmres(c, mK, Rm, VU(i), ls.GM));
        c.vs(mD, VU(i), VP(i));
        hipLaunchKernelGGL(k_ns_sub, grid_of(n), dim3(256), 0, s, n, (const double*)Rc, (const double*)VP(i), VP(i));
        SV_HIP(hipMemcpyAsync(cR, VP(i), n * sizeof(double), hipMemcpyDeviceToDevice, s));
        SV_TRY(cg_schur_solve(L, ls.CG, Gt, mG, mL, cR, cX, cP, cP2, cSP, cGP, s));
        SV_HIP(hipMemcpyAsync(VP(i), cR, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        c.sv(mG, VP(i), VMU(iB));
        hipLaunchKernelGGL(k_ns_sub, grid_of(3 * n), dim3(256), 0, s, 3 * n, (const double*)Rm,
                           (const double*)VMU(iB), VMU(iBB));
        SV_TRY(ns_gmres(c, mK, VMU(iBB), VU(i), ls.GM));
        SV_TRY(launch_spmv(L, kNsd, mK, VU(i), VMU(iBB), s));
        SV_TRY(cpl_apply(L, 0, kNsd, VU(i), VMU(iBB), s));  // NSSOLVER.f:128
        c.ss(mL, VP(i), VMP(iB));
        c.vs(mD, VU(i), VMP(iBB));
        // A(j,k) = MU_j . MU_k + MP_j . MP_k, B(k) = MU_k . Rmi + MP_k . Rci
        // for k = iB, iBB: all pairs in one pass, one read-back
        const int np = 2 * iB + 3;
        hipLaunchKernelGGL(k_ns_gram, dim3(kGramBlocks, np), dim3(256), 0, s, (int)n, iB, (const double*)MU,
                           (const double*)MP, (const double*)Rmi, (const double*)Rci, gpart);
        hipLaunchKernelGGL(k_ns_gram_sum, dim3(2 * np), dim3(64), 0, s, kGramBlocks, (const double*)gpart, gout);
        SV_HIP(hipGetLastError());
        SV_HIP(hipMemcpyAsync(gram.data(), gout, 2 * np * sizeof(double), hipMemcpyDeviceToHost, s));
        SV_HIP(hipStreamSynchronize(s));
        for (int p = 0; p < np; ++p) {
            const int k = p <= iB ? iB : iBB;
            const int j = p <= iB ? p : p - iB - 1;
            const double t = gram[2 * p] + gram[2 * p + 1];
            if (j == 0) {
                B[k - 1] = t;
            } else {
                AM(j, k) = t;
                AM(k, j) = t;
            }
        }
        xB = B;
        if (ge_host(nB, iBB, A, xB)) {
            oldxB = xB;
        } else {
            xB = oldxB;
            if (i > 1) {
                iB -= 2;
                iBB -= 2;
            }
            break;
        }
        double sxb = 0.0;
        for (int k = 0; k < iBB; ++k) sxb += xB[k] * B[k];
        ls.RI.fNorm = ls.RI.iNorm * ls.RI.iNorm - sxb;
        if (ls.RI.fNorm < eps * eps) {
            ls.RI.suc = 1;
            break;
        }
        // Rm = Rmi - xB(1) MU1 - xB(2) MU2 - ..., Rc likewise
        SV_HIP(hipMemcpyAsync(Rm, Rmi, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
        SV_HIP(hipMemcpyAsync(Rc, Rci, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        for (int j = 1; j <= iBB; ++j) {
            c.axpy(Rm, 1.0, -xB[j - 1], VMU(j), 3 * n);
            c.axpy(Rc, 1.0, -xB[j - 1], VMP(j), n);
        }
    }
    if (i > ls.RI.mItr) {
        ls.RI.itr = ls.RI.mItr;
    } else {
        ls.RI.itr = i;
        SV_HIP(hipMemcpyAsync(Rc, Rci, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        for (int j = 1; j <= iBB; ++j) c.axpy(Rc, 1.0, -xB[j - 1], VMP(j), n);
    }
    const double nrc = std::sqrt(c.dot(Rc, Rc, n, &rc));
    SV_TRY(rc);
    ls.Resc = (int)std::lround(100.0 * nrc * nrc / ls.RI.fNorm);
    ls.Resm = 100 - ls.Resc;
    // solution: Rmi = sum xB(2k) U_k, Rci = sum xB(2k-1) P_k
    c.axpy(Rmi, 0.0, xB[1], VU(1), 3 * n);
    c.axpy(Rci, 0.0, xB[0], VP(1), n);
    for (int k = 2; k <= ls.RI.itr; ++k) {
        c.axpy(Rmi, 1.0, xB[2 * k - 1], VU(k), 3 * n);
        c.axpy(Rci, 1.0, xB[2 * k - 2], VP(k), n);
    }
    ls.RI.callD = wall_seconds() - ls.RI.callD;
    ls.RI.dB = 5.0 * std::log(ls.RI.fNorm / ls.RI.dB);
    if (ls.Resc < 0 || ls.Resm < 0) {
        ls.Resc = 0;
        ls.Resm = 0;
        ls.RI.dB = 0.0;
        ls.RI.fNorm = 0.0;
    }
    ls.RI.fNorm = std::sqrt(ls.RI.fNorm);
    hipLaunchKernelGGL(k_ns_merge, grid_of(n), dim3(256), 0, s, (int)n, (const double*)Rmi, (const double*)Rci, Ri);
    SV_HIP(hipGetLastError());
    SV_HIP(hipStreamSynchronize(s));
    return SVFSI_OK;
}

} // namespace svfsi
