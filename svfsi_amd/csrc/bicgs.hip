// bicgs.hip -- BICGSV (svFSILS/BICGS.f:115-180): unpreconditioned BiCGStab
// on the scaled system, FP64, dof 2..4.  Two block SpMVs per iteration; the
// vector updates are single passes, and the reference's five reductions are
// taken as three read-backs: (Rh . V), (T . T, T . S) together, and
// (R . R, R . Rh) together.
#include "internal.hpp"
#include "devutil.hpp"

#include <cmath>

namespace svfsi {

namespace {

constexpr int kDotBlocks = 256;

// block partials of a1 . b1 (and a2 . b2 when a2 is not null)
__global__ __launch_bounds__(256) void k_bi_dot2(size_t n, const double* __restrict__ a1,
                                                 const double* __restrict__ b1, const double* __restrict__ a2,
                                                 const double* __restrict__ b2, double* __restrict__ part)
{
    __shared__ double lds[4];
    double s1 = 0.0, s2 = 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        s1 += a1[q] * b1[q];
        if (a2) s2 += a2[q] * b2[q];
    }
    s1 = block_sum<256>(s1, lds);
    s2 = block_sum<256>(s2, lds);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s1;
        part[gridDim.x + blockIdx.x] = s2;
    }
}

__global__ __launch_bounds__(64) void k_bi_sum2(int nb, const double* __restrict__ part, double* __restrict__ out)
{
    double v = 0.0;
    for (int q = threadIdx.x; q < nb; q += 64) v += part[(size_t)blockIdx.x * nb + q];
    v = wave_sum(v);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// S = R - alpha V
__global__ __launch_bounds__(256) void k_bi_s(size_t n, const double* __restrict__ R, double alpha,
                                              const double* __restrict__ V, double* __restrict__ S)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) S[q] = R[q] - alpha * V[q];
}

// X = X + alpha P + omega S; R = S - omega T
__global__ __launch_bounds__(256) void k_bi_xr(size_t n, double* __restrict__ X, double* __restrict__ R,
                                               const double* __restrict__ P, const double* __restrict__ S,
                                               const double* __restrict__ T, double alpha, double omega)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    X[q] = X[q] + alpha * P[q] + omega * S[q];
    R[q] = S[q] - omega * T[q];
}

// P = R + beta (P - omega V)
__global__ __launch_bounds__(256) void k_bi_p(size_t n, double* __restrict__ P, const double* __restrict__ R,
                                              const double* __restrict__ V, double beta, double omega)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) P[q] = R[q] + beta * (P[q] - omega * V[q]);
}

inline dim3 grid_of(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

} // namespace

int bicgs_solve(Lhs& L, FSILS_subLsType& ls, int dof, const double* K, double* R, hipStream_t s)
{
    if (dof < 2 || dof > 4) {
        set_error("BiCGStab: dof %d not supported (2..4)", dof);
        return SVFSI_ERR_NOTSUP;
    }
    const size_t n = (size_t)L.nNo * dof;
    Work& w = L.ws;
    SV_TRY(w.bicgs.ensure((6 * n + 2 * kDotBlocks + 2) * sizeof(double)));
    double* P = w.bicgs.as<double>();
    double* Rh = P + n;
    double* X = Rh + n;
    double* V = X + n;
    double* S = V + n;
    double* T = S + n;
    double* part = T + n;
    double* out = part + 2 * kDotBlocks;
    double h[2] = {0.0, 0.0};
    int rc = SVFSI_OK;
    auto dots = [&](const double* a1, const double* b1, const double* a2, const double* b2) {
        hipLaunchKernelGGL(k_bi_dot2, dim3(kDotBlocks), dim3(256), 0, s, n, a1, b1, a2, b2, part);
        hipLaunchKernelGGL(k_bi_sum2, dim3(2), dim3(64), 0, s, kDotBlocks, (const double*)part, out);
        if (hipMemcpyAsync(h, out, 2 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = SVFSI_ERR_HIP;
    };

    ls.callD = wall_seconds();
    ls.suc = 0;
    dots(R, R, nullptr, nullptr);
    SV_TRY(rc);
    double err = std::sqrt(h[0]);
    double errO = err;
    ls.iNorm = err;
    const double eps = std::fmax(ls.absTol, ls.relTol * err);
    double rho = err * err;
    SV_HIP(hipMemsetAsync(X, 0, n * sizeof(double), s));
    SV_HIP(hipMemcpyAsync(P, R, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    SV_HIP(hipMemcpyAsync(Rh, R, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    int i;
    for (i = 1; i <= ls.mItr; ++i) {
        if (err < eps) {
            ls.suc = 1;
            break;
        }
        SV_TRY(launch_spmv(L, dof, K, P, V, s));
        dots(Rh, V, nullptr, nullptr);
        SV_TRY(rc);
        const double alpha = rho / h[0];
        hipLaunchKernelGGL(k_bi_s, grid_of(n), dim3(256), 0, s, n, (const double*)R, alpha, (const double*)V, S);
        SV_TRY(launch_spmv(L, dof, K, S, T, s));
        dots(T, T, T, S);
        SV_TRY(rc);
        double omega = std::sqrt(h[0]);
        omega = h[1] / (omega * omega);
        hipLaunchKernelGGL(k_bi_xr, grid_of(n), dim3(256), 0, s, n, X, R, (const double*)P, (const double*)S,
                           (const double*)T, alpha, omega);
        errO = err;
        dots(R, R, R, Rh);
        SV_TRY(rc);
        err = std::sqrt(h[0]);
        const double rhoO = rho;
        rho = h[1];
        const double beta = rho * alpha / (rhoO * omega);
        hipLaunchKernelGGL(k_bi_p, grid_of(n), dim3(256), 0, s, n, P, (const double*)R, (const double*)V, beta, omega);
    }
    SV_HIP(hipMemcpyAsync(R, X, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    SV_HIP(hipGetLastError());
    SV_HIP(hipStreamSynchronize(s));
    ls.itr = i - 1;
    ls.fNorm = err;
    ls.callD = wall_seconds() - ls.callD;
    ls.dB = (errO < 2.220446049250313e-16) ? 0.0 : 10.0 * std::log(err / errO);
    return SVFSI_OK;
}

} // namespace svfsi
