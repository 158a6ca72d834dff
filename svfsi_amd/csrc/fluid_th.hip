// fluid_th.hip -- CONSTRUCT_FLUID with a Taylor-Hood basis (nFs = 2, P2P1)
// on tet10 meshes for gfx950 (svFSI/FLUID.f:40-190 with FLUID3D_M,
// 192-560, and FLUID3D_C, 813-1084, at vmsFlag = .FALSE.; GETTHOODFS,
// svFSI/FS.f:257-346; GNNxx, svFSI/NN.f:1568-1678; THOOD_ValRC,
// FS.f:352-407).
//
// One 128-thread workgroup per element, elements of one colour per launch
// (no two share a node, so the read-modify-write of Val and R is race free
// and the summation order fixed).  The workgroup first evaluates the
// geometry of all 19 Gauss points in parallel (threads 0..14: GNN and GNNxx
// of the tet10 map at the 15 momentum points; 15..18: GNN at the 4
// continuity points; 19: GNN of the corner tet4 map) into LDS.  Then thread
// t < 100 owns the node pair (a, b) = (t % 10, t / 10): at every Gauss
// point it evaluates the point's kinematics (redundantly in every lane,
// which costs a wavefront the same as one lane) and adds the point's terms
// to its 4 x 4 block; threads with b = 0 build node a's residual.  The
// pressure lives on the corner nodes 0..3, so the velocity-pressure columns
// exist for b < 4 and the continuity rows for a < 4.
#include "internal.hpp"
#include "devutil.hpp"
#include "assemble.hpp"

#include <cmath>

// As the other element routines: no multiply-add contraction, the
// reference's association of products.
#pragma clang fp contract(off)

namespace svfsi {

namespace {

constexpr int kNW = 10, kNQ = 4, kNG1 = 15, kNG2 = 4;

__device__ __forceinline__ bool th_iszero(double a)
{
    const double eps = 2.220446049250313e-16;
    const double aa = fabs(a);
    const double nrm = aa > eps ? aa : eps;
    return (aa / nrm) < 10.0 * eps;
}

// GNN (svFSI/NN.f:1469-1565): xXi = sum_a x_a (x) dN_a/dxi, Jac, ks =
// xiX^T xiX, Nx = dN/dxi . xiX, in the reference's operation order.
template <int EN>
__device__ void th_gnn(const double (*xl)[3], const double* Nxi, double (*Nx)[3], double& Jac, double ks[3][3])
{
    double A[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    for (int a = 0; a < EN; ++a)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) A[i][j] += xl[a][i] * Nxi[a * 3 + j];
    const double J = A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0]
                   + A[0][2] * A[1][0] * A[2][1] - A[0][0] * A[1][2] * A[2][1]
                   - A[0][1] * A[1][0] * A[2][2] - A[0][2] * A[1][1] * A[2][0];
    double B[3][3];
    B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / J;
    B[0][1] = (A[2][1] * A[0][2] - A[2][2] * A[0][1]) / J;
    B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / J;
    B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / J;
    B[1][1] = (A[2][2] * A[0][0] - A[2][0] * A[0][2]) / J;
    B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / J;
    B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / J;
    B[2][1] = (A[2][0] * A[0][1] - A[2][1] * A[0][0]) / J;
    B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / J;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ks[i][j] = B[0][i] * B[0][j] + B[1][i] * B[1][j] + B[2][i] * B[2][j];
    for (int a = 0; a < EN; ++a)
        for (int j = 0; j < 3; ++j)
            Nx[a][j] = Nxi[a * 3 + 0] * B[0][j] + Nxi[a * 3 + 1] * B[1][j] + Nxi[a * 3 + 2] * B[2][j];
    Jac = J;
}

// GETGNNxx, TET10 (svFSI/NN.f:1285-1295): constant d2N/dxi_i dxi_j, order
// (11, 22, 33, 12, 23, 13).
// (a table in constant memory: indexed by a run-time node, a local array
// would be materialised in registers)
__constant__ double kTet10Xixi[10][6] = {{4, 0, 0, 0, 0, 0}, {0, 4, 0, 0, 0, 0}, {0, 0, 4, 0, 0, 0},
                                         {4, 4, 4, 4, 4, 4}, {0, 0, 0, 4, 0, 0}, {0, 0, 0, 0, 4, 0},
                                         {0, 0, 0, 0, 0, 4}, {-8, 0, 0, -4, 0, -4}, {0, -8, 0, -4, -4, 0},
                                         {0, 0, -8, 0, -4, -4}};

__device__ __forceinline__ double tet10_xixi(int a, int i)
{
    return kTet10Xixi[a][i];
}

// GNNxx (svFSI/NN.f:1568-1678), 3-D: the 6 x 6 chain-rule system K Nxx(a) =
// Nxi2(a) - Nx(a) . xXi2 for the physical second derivatives (order xx, yy,
// zz, xy, yz, xz), solved by LU with partial pivoting (the algorithm of
// the DGESV call at NN.f:1672) for the ten right-hand sides.
__device__ void th_gnnxx_lu(const double (*xl)[3], const double* Nxi, double* LU, int* pivOut, double* X2)
{
    double xXi[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double xXi2[3][6];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 6; ++i) xXi2[k][i] = 0.0;
    for (int a = 0; a < kNW; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
#pragma unroll
            for (int i = 0; i < 3; ++i) xXi[k][i] += xl[a][k] * Nxi[a * 3 + i];
#pragma unroll
            for (int i = 0; i < 6; ++i) xXi2[k][i] += xl[a][k] * tet10_xixi(a, i);
        }
    double K[6][6];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        K[i][0] = xXi[0][i] * xXi[0][i];
        K[i][1] = xXi[1][i] * xXi[1][i];
        K[i][2] = xXi[2][i] * xXi[2][i];
        K[i][3] = 2.0 * xXi[0][i] * xXi[1][i];
        K[i][4] = 2.0 * xXi[1][i] * xXi[2][i];
        K[i][5] = 2.0 * xXi[0][i] * xXi[2][i];
    }
    constexpr int pr[3][2] = {{0, 1}, {1, 2}, {0, 2}};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int i = pr[r][0], j = pr[r][1];
        K[3 + r][0] = xXi[0][i] * xXi[0][j];
        K[3 + r][1] = xXi[1][i] * xXi[1][j];
        K[3 + r][2] = xXi[2][i] * xXi[2][j];
        K[3 + r][3] = xXi[0][i] * xXi[1][j] + xXi[0][j] * xXi[1][i];
        K[3 + r][4] = xXi[1][i] * xXi[2][j] + xXi[1][j] * xXi[2][i];
        K[3 + r][5] = xXi[0][i] * xXi[2][j] + xXi[0][j] * xXi[2][i];
    }
    // LU with partial pivoting, row swaps over compile-time indices (a
    // lane-dependent index would move the matrix to scratch memory)
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int p = k;
        double best = fabs(K[k][k]);
#pragma unroll
        for (int i = k + 1; i < 6; ++i)
            if (fabs(K[i][k]) > best) {
                best = fabs(K[i][k]);
                p = i;
            }
        pivOut[k] = p;
#pragma unroll
        for (int i = k + 1; i < 6; ++i)
            if (i == p)
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    const double tt = K[k][j];
                    K[k][j] = K[i][j];
                    K[i][j] = tt;
                }
#pragma unroll
        for (int i = k + 1; i < 6; ++i) {
            K[i][k] /= K[k][k];
#pragma unroll
            for (int j = k + 1; j < 6; ++j) K[i][j] -= K[i][k] * K[k][j];
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) LU[i * 6 + j] = K[i][j];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 6; ++i) X2[k * 6 + i] = xXi2[k][i];
}

// one right-hand side of GNNxx: shape function a, with the point's LU
// factors, pivots and x_xixi from th_gnnxx_lu
__device__ void th_gnnxx_rhs(int a, const double* Nxa, const double* LU, const int* piv, const double* X2,
                             double* out)
{
    double b[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) b[i] = tet10_xixi(a, i) - Nxa[0] * X2[i] - Nxa[1] * X2[6 + i] - Nxa[2] * X2[12 + i];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int pk = piv[k];
#pragma unroll
        for (int i = k + 1; i < 6; ++i)
            if (i == pk) {
                const double tt = b[k];
                b[k] = b[i];
                b[i] = tt;
            }
    }
#pragma unroll
    for (int i = 1; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) b[i] -= LU[i * 6 + j] * b[j];
#pragma unroll
    for (int i = 5; i >= 0; --i) {
#pragma unroll
        for (int j = i + 1; j < 6; ++j) b[i] -= LU[i * 6 + j] * b[j];
        b[i] /= LU[i * 6 + i];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) out[i] = b[i];
}

// GETVISCOSITY (svFSI/FLUID.f:1690-1739) at the shear rate gam, then
// mu_g = (d mu / d gamma) / gamma (FLUID.f:375-381).
__device__ void th_visc(const AsmParams& P, double mu_i, double gam, double& mu, double& mu_g)
{
    mu = mu_i;
    mu_g = 0.0;
    if (P.viscType != 697 && P.viscType != 696) return;
    double mux;
    if (P.viscType == 697) {
        double T1 = 1.0 + pow(P.lam * gam, P.va);
        double T2 = pow(T1, (P.vn - 1.0) / P.va);
        mu = mu_i + (P.mu_o - mu_i) * T2;
        T1 = T2 / T1;
        T2 = pow(P.lam, P.va) * pow(gam, P.va - 1.0) * T1;
        mux = (P.mu_o - mu_i) * (P.vn - 1.0) * T2;
    } else {
        double mu_o = P.mu_o;
        if (gam < P.lam) {
            mu_o = mu_o / sqrt(P.lam);
            gam = P.lam;
        } else {
            mu_o = mu_o / sqrt(gam);
        }
        mu = (mu_i + mu_o) * (mu_i + mu_o);
        mux = 2.0 * mu_o * (mu_o + mu_i) / gam;
    }
    mu_g = th_iszero(gam) ? 0.0 : mux / gam;
}

__global__ __launch_bounds__(128) void k_asm_fluid_th(int begin, const int* __restrict__ ien,
                                                      const int* __restrict__ elemId, const int* __restrict__ emap,
                                                      const double* __restrict__ x, int tDof,
                                                      const double* __restrict__ Ag, const double* __restrict__ Yg,
                                                      const double* __restrict__ muEl, AsmParams P,
                                                      const double* __restrict__ rule, double* __restrict__ val,
                                                      double* __restrict__ R, double* __restrict__ ke,
                                                      double* __restrict__ re, const int* __restrict__ gSlot,
                                                      const int* __restrict__ rSlot)
{
    const int p = begin + blockIdx.x;
    const int t = threadIdx.x;
    __shared__ double sx[kNW][3], sa[kNW][3], sy[kNW][4];
    __shared__ double sNwx[kNG1 + kNG2][kNW][3];
    __shared__ double sNwxx[kNG1][kNW][6];
    __shared__ double sW[kNG1 + kNG2];
    __shared__ double sKs[kNG1 + kNG2][3][3];
    __shared__ double sNqx[kNQ][3];
    __shared__ int sNode[kNW];
    __shared__ double sLU[kNG1][36], sX2[kNG1][18];
    __shared__ int sPiv[kNG1][6];
    if (t < kNW) {
        const int n = ien[(size_t)p * kNW + t];
        sNode[t] = n;
        for (int i = 0; i < 3; ++i) sx[t][i] = x[(size_t)n * 3 + i];
        for (int i = 0; i < 3; ++i) sa[t][i] = Ag[(size_t)n * tDof + i];
        for (int i = 0; i < 4; ++i) sy[t][i] = Yg[(size_t)n * tDof + i];
    }
    __syncthreads();
    // Gauss-point geometry (GETTHOODFS / GNN / GNNxx bookkeeping of
    // CONSTRUCT_FLUID, FLUID.f:103-160)
    if (t < kNG1) {
        // momentum point t: the tet10 map gives Nwx, Jac and ksix, GNNxx Nwxx
        const double* r = rule + (size_t)t * kThRec;
        double ks[3][3], J;
        th_gnn<kNW>(sx, r + 11, sNwx[t], J, ks);
        th_gnnxx_lu(sx, r + 11, sLU[t], sPiv[t], sX2[t]);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) sKs[t][i][j] = ks[i][j];
        sW[t] = r[0] * J;
    } else if (t < kNG1 + kNG2 + 1) {
        // continuity point g = t - 15: the tet10 map, except that at the
        // first point the corner tet4 map's Jac and ksix replace it; t = 19:
        // the tet4 map's Nqx (constant)
        constexpr double Nqxi[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, -1, -1, -1};
        double ks[3][3], J;
        if (t < kNG1 + kNG2) {
            const int g = t - kNG1;
            const double* r = rule + (size_t)t * kThRec;
            th_gnn<kNW>(sx, r + 11, sNwx[t], J, ks);
            if (g == 0) {
                double Nq[kNQ][3];
                th_gnn<kNQ>(sx, Nqxi, Nq, J, ks);
            }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) sKs[t][i][j] = ks[i][j];
            sW[t] = r[0] * J;
        } else {
            double Nq[kNQ][3];
            th_gnn<kNQ>(sx, Nqxi, Nq, J, ks);
            for (int a = 0; a < kNQ; ++a)
                for (int i = 0; i < 3; ++i) sNqx[a][i] = Nq[a][i];
        }
    }
    __syncthreads();
    // GNNxx right-hand sides: (point, shape function) pairs over the threads
    for (int q = t; q < kNG1 * kNW; q += 128) {
        const int g = q / kNW, a = q - (q / kNW) * kNW;
        th_gnnxx_rhs(a, sNwx[g][a], sLU[g], sPiv[g], sX2[g], sNwxx[g][a]);
    }
    __syncthreads();

    const bool own = t < kNW * kNW;
    const int a = t % kNW, b = own ? t / kNW : 0;
    const double rho = P.rho;
    const double T1c = P.af * P.gam * P.dt;
    const double amd = P.am / T1c;
    const double mu_i = muEl ? muEl[elemId[p]] : P.mu;
    double k[16], lr[4];
    for (int q = 0; q < 16; ++q) k[q] = 0.0;
    for (int q = 0; q < 4; ++q) lr[q] = 0.0;

    // ---- integration 1: FLUID3D_M at the 15 points of the tet10 rule.
    // Point kinematics first, one thread per point (the same expressions in
    // the same order as a per-lane evaluation would use), into LDS
    __shared__ double sU[kNG1][3], sUd[kNG1][3], sUx[kNG1][3][3], sEs[kNG1][3][3], sD2u2[kNG1][3];
    __shared__ double sPp[kNG1], sMu[kNG1], sMuG[kNG1], sMuX[kNG1][3], sTauM[kNG1], sUp[kNG1][3];
    __shared__ double sRM[kNG1][3][3], sRV[kNG1][3];
    if (t < kNG1) {
        const int g = t;
        const double* r = rule + (size_t)g * kThRec;
        const double* Nw = r + 1;
        const double* Nq = r + 41;
        // velocity, its gradient and second derivatives, inertia (FLUID.f:228-291)
        double ud[3], u[3], ux[3][3], uxx[3][3][3];
        for (int c = 0; c < 3; ++c) {
            ud[c] = -P.fb[c];
            u[c] = 0.0;
            for (int i = 0; i < 3; ++i) {
                ux[i][c] = 0.0;
                for (int j = 0; j < 3; ++j) uxx[i][c][j] = 0.0;
            }
        }
        // node loop kept rolled: unrolled, its loads filled the register
        // file (one wave per SIMD); rolled, the kernel runs two
#pragma unroll 1
        for (int n = 0; n < kNW; ++n) {
            for (int c = 0; c < 3; ++c) {
                ud[c] += Nw[n] * (sa[n][c] - 0.0);
                u[c] += Nw[n] * sy[n][c];
                for (int i = 0; i < 3; ++i) ux[i][c] += sNwx[g][n][i] * sy[n][c];
                uxx[0][c][0] += sNwxx[g][n][0] * sy[n][c];
                uxx[1][c][1] += sNwxx[g][n][1] * sy[n][c];
                uxx[2][c][2] += sNwxx[g][n][2] * sy[n][c];
                uxx[1][c][0] += sNwxx[g][n][3] * sy[n][c];
                uxx[2][c][1] += sNwxx[g][n][4] * sy[n][c];
                uxx[0][c][2] += sNwxx[g][n][5] * sy[n][c];
            }
        }
        const double divU = ux[0][0] + ux[1][1] + ux[2][2];
        double d2u2[3];
        for (int c = 0; c < 3; ++c) {
            uxx[0][c][1] = uxx[1][c][0];
            uxx[1][c][2] = uxx[2][c][1];
            uxx[2][c][0] = uxx[0][c][2];
            d2u2[c] = uxx[0][c][0] + uxx[1][c][1] + uxx[2][c][2];
        }
        (void)divU;
        // pressure on the corner nodes (FLUID.f:293-301)
        double pp = 0.0, px[3] = {0.0, 0.0, 0.0};
        for (int n = 0; n < kNQ; ++n) {
            pp += Nq[n] * sy[n][3];
            for (int i = 0; i < 3; ++i) px[i] += sNqx[n][i] * sy[n][3];
        }
        // strain rate, its gradient, shear rate and viscosity (FLUID.f:312-381)
        double es[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) es[i][j] = ux[i][j] + ux[j][i];
        double mu_x[3];
        for (int m = 0; m < 3; ++m) {
            double ex[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) ex[i][j] = uxx[i][j][m] + uxx[j][i][m];
            mu_x[m] = (ex[0][0] * es[0][0] + ex[1][1] * es[1][1] + ex[2][2] * es[2][2]) * 0.5
                    + ex[1][0] * es[1][0] + ex[2][1] * es[2][1] + ex[0][2] * es[0][2];
        }
        double gam = 0.0;
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) gam += es[i][j] * es[i][j];
        gam = sqrt(0.5 * gam);
        double mu, mu_g;
        th_visc(P, mu_i, gam, mu, mu_g);
        for (int m = 0; m < 3; ++m) mu_x[m] = mu_g * mu_x[m];
        // stabilisation: tau_M and the fine-scale velocity (FLUID.f:383-402);
        // without VMS tau_C = tau_B = 0, ua = u, pa = p (FLUID.f:421-426)
        const double (*Ks)[3] = sKs[g];
        const double kT = 4.0 * (1.0 / P.dt) * (1.0 / P.dt);
        double kU = 0.0, kS = 0.0;
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                kU += u[i] * u[j] * Ks[i][j];
                kS += Ks[i][j] * Ks[i][j];
            }
        kS = 36.0 * kS * (mu / rho) * (mu / rho);
        const double tauM = 1.0 / (rho * sqrt(kT + kU + kS));
        double up[3];
        for (int c = 0; c < 3; ++c) {
            const double rV = ud[c] + u[0] * ux[0][c] + u[1] * ux[1][c] + u[2] * ux[2][c];
            const double rS = mu_x[0] * es[0][c] + mu_x[1] * es[1][c] + mu_x[2] * es[2][c] + mu * d2u2[c];
            up[c] = -tauM * (rho * rV + px[c] - rS);
        }
        const double tauC = 0.0, tauB = 0.0;
        double rM[3][3], rV[3];
        for (int c = 0; c < 3; ++c) {
            const double rv = tauB * (up[0] * ux[0][c] + up[1] * ux[1][c] + up[2] * ux[2][c]);
            for (int i = 0; i < 3; ++i)
                rM[i][c] = mu * es[i][c] - rho * up[c] * u[i] + rv * up[i] - (i == c ? pp : 0.0);
        }
        for (int c = 0; c < 3; ++c) rV[c] = ud[c] + u[0] * ux[0][c] + u[1] * ux[1][c] + u[2] * ux[2][c];
        for (int c = 0; c < 3; ++c) {
            sU[g][c] = u[c];
            sUd[g][c] = ud[c];
            sD2u2[g][c] = d2u2[c];
            sMuX[g][c] = mu_x[c];
            sUp[g][c] = up[c];
            sRV[g][c] = rV[c];
            for (int i = 0; i < 3; ++i) {
                sUx[g][i][c] = ux[i][c];
                sEs[g][i][c] = es[i][c];
                sRM[g][i][c] = rM[i][c];
            }
        }
        sPp[g] = pp;
        sMu[g] = mu;
        sMuG[g] = mu_g;
        sTauM[g] = tauM;
        (void)Nw;
    }
    __syncthreads();
    for (int g = 0; g < kNG1; ++g) {
        const double* r = rule + (size_t)g * kThRec;
        const double* Nw = r + 1;
        const double* Nq = r + 41;
        const double w = sW[g];
        const double wl = w * T1c, wr = w * rho;
        const double* u = sU[g];
        const double* d2u2 = sD2u2[g];
        const double* mu_x = sMuX[g];
        const double* up = sUp[g];
        const double* rV = sRV[g];
        const double (*es)[3] = sEs[g];
        const double (*rM)[3] = sRM[g];
        const double mu = sMu[g], mu_g = sMuG[g], tauM = sTauM[g];
        const double tauC = 0.0, tauB = 0.0;
        if (!own) continue;
        const double* Xa = sNwx[g][a];
        const double* Xb = sNwx[g][b];
        // residual of node a (FLUID.f:445-454)
        if (b == 0)
            for (int c = 0; c < 3; ++c)
                lr[c] += wr * Nw[a] * rV[c] + w * (Xa[0] * rM[0][c] + Xa[1] * rM[1][c] + Xa[2] * rM[2][c]);
        // per-node quantities of a and b (FLUID.f:456-479)
        const double uNxa = u[0] * Xa[0] + u[1] * Xa[1] + u[2] * Xa[2];
        const double uNxb = u[0] * Xb[0] + u[1] * Xb[1] + u[2] * Xb[2];
        const double upNxa = up[0] * Xa[0] + up[1] * Xa[1] + up[2] * Xa[2];
        const double upNxb = up[0] * Xb[0] + up[1] * Xb[1] + up[2] * Xb[2];
        const double uaNxa = uNxa;
        double esNxa[3], esNxb[3];
        for (int c = 0; c < 3; ++c) {
            esNxa[c] = es[0][c] * Xa[0] + es[1][c] * Xa[1] + es[2][c] * Xa[2];
            esNxb[c] = es[0][c] * Xb[0] + es[1][c] * Xb[1] + es[2][c] * Xb[2];
        }
        const double* Nxxb = sNwxx[g][b];
        const double T1b = -rho * uNxb + mu * (Nxxb[0] + Nxxb[1] + Nxxb[2]) + mu_x[0] * Xb[0] + mu_x[1] * Xb[1]
                         + mu_x[2] * Xb[2];
        // velocity-velocity block (FLUID.f:482-545)
        const double NxNx = Xa[0] * Xb[0] + Xa[1] * Xb[1] + Xa[2] * Xb[2];
        const double T1 = mu * NxNx + rho * amd * Nw[b] * (Nw[a] + rho * tauM * uaNxa)
                        + rho * Nw[a] * (uNxb + upNxb) + tauB * upNxa * upNxb;
        for (int rr = 0; rr < 3; ++rr)
            for (int c = 0; c < 3; ++c) {
                // updu(c, rr, b) (FLUID.f:467-478)
                double updu = mu_x[c] * Xb[rr] + d2u2[rr] * mu_g * esNxb[c];
                if (c == rr) updu = updu + T1b;
                const double T2 = mu * (Xa[c] * Xb[rr]) + tauC * (Xa[rr] * Xb[c]) + esNxa[rr] * mu_g * esNxb[c]
                                - rho * tauM * uaNxa * updu;
                k[rr * 4 + c] += wl * (rr == c ? T2 + T1 : T2);
            }
        // velocity-pressure block, pressure on the corner nodes (FLUID.f:547-557)
        if (b < kNQ) {
            const double T1p = rho * tauM * uaNxa;
            for (int rr = 0; rr < 3; ++rr) k[rr * 4 + 3] -= wl * (Xa[rr] * Nq[b] - sNqx[b][rr] * T1p);
        }
    }

    // ---- integration 2: FLUID3D_C at the 4 points of the tet4 rule; Nwxx
    // keeps its last value, and without VMS tau_M = u' = updu = 0
    // (FLUID.f:991-1043), so only div u and the basis enter
    for (int g = 0; g < kNG2; ++g) {
        const int gg = kNG1 + g;
        const double* r = rule + (size_t)gg * kThRec;
        const double* Nw = r + 1;
        const double* Nq = r + 41;
        const double w = sW[gg];
        const double wl = w * T1c;
        double ux00 = 0.0, ux11 = 0.0, ux22 = 0.0;
        for (int n = 0; n < kNW; ++n) {
            ux00 += sNwx[gg][n][0] * sy[n][0];
            ux11 += sNwx[gg][n][1] * sy[n][1];
            ux22 += sNwx[gg][n][2] * sy[n][2];
        }
        const double divU = ux00 + ux11 + ux22;
        if (!own || a >= kNQ) continue;
        if (b == 0) lr[3] += w * (Nq[a] * divU - 0.0);
        (void)Nw;
        const double* Xb = sNwx[gg][b];
        for (int c = 0; c < 3; ++c) k[12 + c] += wl * (Nq[a] * Xb[c]);
    }

    if (!own) return;
    if (ke) {
        // gather assembly: the element block at its gather position (the
        // sums into Val follow in element-id order)
        double* kb = ke + (size_t)gSlot[((size_t)p * kNW + a) * kNW + b] * 16;
        for (int q = 0; q < 16; ++q) kb[q] = k[q];
        if (b == 0) {
            double* rb = re + (size_t)rSlot[(size_t)p * kNW + a] * 4;
            for (int c = 0; c < 4; ++c) rb[c] = lr[c];
        }
        return;
    }
    // DOASSEM through the scatter map (colours keep the updates disjoint)
    double* vb = val + (size_t)emap[((size_t)p * kNW + a) * kNW + b] * 16;
    for (int q = 0; q < 16; ++q) vb[q] += k[q];
    if (b == 0) {
        double* rr = R + (size_t)sNode[a] * 4;
        for (int c = 0; c < 4; ++c) rr[c] += lr[c];
    }
}

// THOOD_ValRC (svFSI/FS.f:352-407): continuity residual 0 and a unit
// pressure diagonal in the rows of the mid-edge nodes.
__global__ __launch_bounds__(256) void k_thood_valrc(int nMid, const int* __restrict__ mid,
                                                     const int* __restrict__ rowPtr, const int* __restrict__ colPtr,
                                                     double* __restrict__ val, double* __restrict__ R)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nMid) return;
    const int n = mid[q];
    R[(size_t)n * 4 + 3] = 0.0;
    for (int k = rowPtr[n]; k < rowPtr[n + 1]; ++k) val[(size_t)k * 16 + 15] = colPtr[k] == n ? 1.0 : 0.0;
}

} // namespace

int launch_asm_fluid_th(Mesh& M, const AsmParams& P, int tDof, const double* Ag, const double* Yg,
                        const double* muEl, double* val, double* R, hipStream_t s)
{
    if (M.eNoN != 10) {
        set_error("Taylor-Hood FLUID needs a tet10 mesh");
        return SVFSI_ERR_ARG;
    }
    if (!M.ruleTH.p) {
        std::vector<double> tab;
        make_th_rule(tab);
        SV_TRY(M.ruleTH.ensure(tab.size() * sizeof(double)));
        SV_HIP(hipMemcpy(M.ruleTH.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    // Gather (default): one launch over all elements stores the element
    // blocks, launch_gather sums them in element-id order; SVFSI_ASM_GATHER=0
    // keeps the coloured read-modify-write, one launch per colour.
    const char* ge = getenv("SVFSI_ASM_GATHER");
    const bool gather = !(ge && ge[0] == '0');
    if (gather) {
        SV_TRY(M.ke.ensure((size_t)M.nEl * 100 * 16 * sizeof(double)));
        SV_TRY(M.re.ensure((size_t)M.nEl * 10 * 4 * sizeof(double)));
        hipLaunchKernelGGL(k_asm_fluid_th, dim3(M.nEl), dim3(128), 0, s, 0, M.ien.as<int>(), M.elemId.as<int>(),
                           M.emap.as<int>(), M.x.as<double>(), tDof, Ag, Yg, muEl, P, M.ruleTH.as<double>(), val, R,
                           M.ke.as<double>(), M.re.as<double>(), M.gSlot.as<int>(), M.rSlot.as<int>());
        SV_HIP(hipGetLastError());
        SV_TRY(launch_gather(M, 4, val, R, s));
    } else {
        const int nC = (int)M.colorOff.size() - 1;
        for (int c = 0; c < nC; ++c) {
            const int begin = M.colorOff[c], count = M.colorOff[c + 1] - begin;
            if (count == 0) continue;
            hipLaunchKernelGGL(k_asm_fluid_th, dim3(count), dim3(128), 0, s, begin, M.ien.as<int>(),
                               M.elemId.as<int>(), M.emap.as<int>(), M.x.as<double>(), tDof, Ag, Yg, muEl, P,
                               M.ruleTH.as<double>(), val, R, nullptr, nullptr, nullptr, nullptr);
        }
        SV_HIP(hipGetLastError());
    }
    if (M.nMid > 0) {
        hipLaunchKernelGGL(k_thood_valrc, dim3((M.nMid + 255) / 256), dim3(256), 0, s, M.nMid, M.midNodes.as<int>(),
                           M.lhs->rowPtr.as<int>(), M.lhs->colPtr.as<int>(), val, R);
        SV_HIP(hipGetLastError());
    }
    return SVFSI_OK;
}

} // namespace svfsi
