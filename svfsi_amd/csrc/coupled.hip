// coupled.hip -- coupled Neumann (resistance) faces of FSILS_SOLVE: the
// rank-one terms res v v^T that ADDBCMUL adds to every product of GMRESV
// and NSSOLVER, and the BCPRE inverse (I + res v v^T)^-1 that the NS
// solver's GMRES and Schur-complement CG apply (svFSILS/ADDBCMUL.f:53-114,
// GMRES.f:393-429, NSSOLVER.f:307-338, PRECOND.f:133-140).
//
// A face is a few hundred to a few thousand nodes, so each face is one
// workgroup: the dot v . X over its nodes (fixed-order tree reduction in
// LDS) and then the update of the same nodes.  Faces without shared nodes
// run as one launch and in place; faces that share nodes run one launch per
// face in the reference's order, reading a copy of X when X is Y.
#include "internal.hpp"
#include "devutil.hpp"

#include <vector>

namespace svfsi {

namespace {

constexpr int kCplThreads = 256;

__global__ __launch_bounds__(kCplThreads) void k_cpl_apply(int f0, int nC, int op, int vd, int dof,
                                                           const int* __restrict__ off, const int* __restrict__ glob,
                                                           const int* __restrict__ fdof, const double* __restrict__ vm,
                                                           const double* __restrict__ coef, const double* X, double* Y,
                                                           const int* stop)
{
    if (stop && *stop) return;
    const int f = f0 + blockIdx.x;
    const int b = off[f], e = off[f + 1];
    const int nsd = fdof[f] < vd ? fdof[f] : vd;
    __shared__ double red[kCplThreads];
    double acc = 0.0;
    for (int t = b + (int)threadIdx.x; t < e; t += kCplThreads) {
        const size_t g = (size_t)glob[t] * vd;
        for (int i = 0; i < nsd; ++i) acc += vm[(size_t)t * dof + i] * X[g + i];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kCplThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    // every read of X is behind the barriers: Y may be X
    const double S = coef[op * nC + f] * red[0];
    for (int t = b + (int)threadIdx.x; t < e; t += kCplThreads) {
        const size_t g = (size_t)glob[t] * vd;
        for (int i = 0; i < nsd; ++i) Y[g + i] = Y[g + i] + vm[(size_t)t * dof + i] * S;
    }
}

// valM = val W on the face's own components (PRECOND.f:133-140)
__global__ __launch_bounds__(256) void k_cpl_scale(int nC, int T, int dof, const int* __restrict__ off,
                                                   const int* __restrict__ glob, const int* __restrict__ fdof,
                                                   const double* __restrict__ val, const double* __restrict__ W,
                                                   double* __restrict__ vm)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= T) return;
    int f = 0;
    while (f + 1 < nC && off[f + 1] <= t) ++f;
    const int nv = fdof[f] < dof ? fdof[f] : dof;
    const size_t g = (size_t)glob[t] * dof;
    for (int i = 0; i < dof; ++i) vm[(size_t)t * dof + i] = i < nv ? val[(size_t)t * dof + i] * W[g + i] : 0.0;
}

// BCPRE: nS = |valM|^2 over dof - 1 components; coef[nC+f] = -res/(1 + res nS)
__global__ __launch_bounds__(kCplThreads) void k_cpl_coef(int nC, int dof, const int* __restrict__ off,
                                                          const int* __restrict__ fdof, const double* __restrict__ vm,
                                                          double* __restrict__ coef)
{
    const int f = blockIdx.x;
    const int ns = fdof[f] < dof - 1 ? fdof[f] : dof - 1;
    __shared__ double red[kCplThreads];
    double acc = 0.0;
    for (int t = off[f] + (int)threadIdx.x; t < off[f + 1]; t += kCplThreads)
        for (int i = 0; i < ns; ++i) {
            const double v = vm[(size_t)t * dof + i];
            acc += v * v;
        }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kCplThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double res = coef[f];
        coef[nC + f] = -res / (1.0 + res * red[0]);
    }
}

} // namespace

int cpl_setup(Lhs& L, int dof, const int* incL, const double* res, hipStream_t s)
{
    CplSet& c = L.cpl;
    c.active = false;
    bool anyNeu = false;
    for (const Face& F : L.faces)
        if (F.foC && F.bGrp == BC_TYPE_Neu) anyNeu = true;
    if (!anyNeu) return SVFSI_OK;
    if (!res) {
        set_error("FSILS: res is required for Neu surfaces (SOLVE.f:74-77)");
        return SVFSI_ERR_ARG;
    }
    std::vector<int> ids;
    for (int f = 0; f < L.nFaces; ++f) {
        const Face& F = L.faces[f];
        const bool inc = !incL || incL[f] != 0;
        if (F.foC && inc && F.bGrp == BC_TYPE_Neu && res[f] != 0.0 && F.nNo > 0) ids.push_back(f);
    }
    if (ids.empty()) return SVFSI_OK;
    if (ids != c.ids || c.keyDof != dof || c.keyVersion != L.facesVersion) {
        const int nC = (int)ids.size();
        std::vector<int> off(nC + 1, 0), fd(nC);
        for (int k = 0; k < nC; ++k) {
            off[k + 1] = off[k] + L.faces[ids[k]].nNo;
            fd[k] = L.faces[ids[k]].dof;
        }
        const int T = off[nC];
        std::vector<int> glob(T);
        std::vector<double> val((size_t)T * dof, 0.0);
        std::vector<char> seen(L.nNo, 0);
        bool disjoint = true;
        for (int k = 0; k < nC; ++k) {
            const Face& F = L.faces[ids[k]];
            const int nv = F.dof < dof ? F.dof : dof;
            for (int a = 0; a < F.nNo; ++a) {
                const int t = off[k] + a;
                glob[t] = F.glob[a];
                if (seen[F.glob[a]]) disjoint = false;
                for (int i = 0; i < nv; ++i) val[(size_t)t * dof + i] = F.val[(size_t)a * F.dof + i];
            }
            for (int a = 0; a < F.nNo; ++a) seen[F.glob[a]] = 1;
        }
        SV_TRY(c.off.ensure((nC + 1) * sizeof(int)));
        SV_TRY(c.fdof.ensure(nC * sizeof(int)));
        SV_TRY(c.glob.ensure((size_t)T * sizeof(int)));
        SV_TRY(c.val.ensure((size_t)T * dof * sizeof(double)));
        SV_TRY(c.vm.ensure((size_t)T * dof * sizeof(double)));
        SV_TRY(c.coef.ensure(2 * nC * sizeof(double)));
        SV_TRY(c.tmp.ensure((size_t)L.nNo * dof * sizeof(double)));
        SV_HIP(hipMemcpy(c.off.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
        SV_HIP(hipMemcpy(c.fdof.p, fd.data(), fd.size() * sizeof(int), hipMemcpyHostToDevice));
        SV_HIP(hipMemcpy(c.glob.p, glob.data(), glob.size() * sizeof(int), hipMemcpyHostToDevice));
        SV_HIP(hipMemcpy(c.val.p, val.data(), val.size() * sizeof(double), hipMemcpyHostToDevice));
        c.ids = ids;
        c.keyDof = dof;
        c.keyVersion = L.facesVersion;
        c.nC = nC;
        c.T = T;
        c.dof = dof;
        c.disjoint = disjoint;
        c.gen++;
    }
    c.hres.resize(c.nC);
    for (int k = 0; k < c.nC; ++k) c.hres[k] = res[c.ids[k]];
    SV_HIP(hipMemcpyAsync(c.coef.p, c.hres.data(), c.nC * sizeof(double), hipMemcpyHostToDevice, s));
    c.active = true;
    return SVFSI_OK;
}

int cpl_scale(Lhs& L, const double* W, hipStream_t s)
{
    CplSet& c = L.cpl;
    if (!c.active) return SVFSI_OK;
    hipLaunchKernelGGL(k_cpl_scale, dim3((c.T + 255) / 256), dim3(256), 0, s, c.nC, c.T, c.dof, c.off.as<int>(),
                       c.glob.as<int>(), c.fdof.as<int>(), c.val.as<double>(), W, c.vm.as<double>());
    hipLaunchKernelGGL(k_cpl_coef, dim3(c.nC), dim3(kCplThreads), 0, s, c.nC, c.dof, c.off.as<int>(),
                       c.fdof.as<int>(), c.vm.as<double>(), c.coef.as<double>());
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

int cpl_apply(Lhs& L, int op, int vd, const double* X, double* Y, hipStream_t s, const int* stop)
{
    CplSet& c = L.cpl;
    if (!c.active) return SVFSI_OK;
    if (vd > c.dof) {
        set_error("coupled faces: %d components per node, the solve has %d", vd, c.dof);
        return SVFSI_ERR_ARG;
    }
    if (c.disjoint) {
        hipLaunchKernelGGL(k_cpl_apply, dim3(c.nC), dim3(kCplThreads), 0, s, 0, c.nC, op, vd, c.dof,
                           c.off.as<int>(), c.glob.as<int>(), c.fdof.as<int>(), c.vm.as<double>(),
                           c.coef.as<double>(), X, Y, stop);
    } else {
        const double* src = X;
        if (X == Y) {
            SV_HIP(hipMemcpyAsync(c.tmp.p, X, (size_t)L.nNo * vd * sizeof(double), hipMemcpyDeviceToDevice, s));
            src = c.tmp.as<double>();
        }
        for (int f = 0; f < c.nC; ++f)
            hipLaunchKernelGGL(k_cpl_apply, dim3(1), dim3(kCplThreads), 0, s, f, c.nC, op, vd, c.dof,
                               c.off.as<int>(), c.glob.as<int>(), c.fdof.as<int>(), c.vm.as<double>(),
                               c.coef.as<double>(), src, Y, stop);
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

} // namespace svfsi
