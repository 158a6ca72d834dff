// assemble.hip -- element tangent/residual construction and DOASSEM scatter
// for gfx950.
//
// Elements are processed colour by colour (no two elements of a colour share
// a node), so every block of Val and every row of R receives at most one
// contribution per launch and the summation order is fixed: the assembly is
// deterministic bit for bit.  Within a colour, one thread owns one (a,b)
// node pair of one element: it recomputes the element geometry and the
// Gauss-point kinematics (cheap, register resident) and accumulates its
// dof x dof block over the Gauss points exactly as the reference routine
// does for that block, then adds it to Val through the precomputed scatter
// map (the bisection of DOASSEM, svFSI/LHSA.f:282-292, done once at mesh
// creation).  The thread with b = 0 also builds the residual rows of node a.
#include "internal.hpp"
#include "devutil.hpp"
#include "assemble.hpp"

#include <cmath>

// The element routines are evaluated without contracting multiply-adds into
// FMAs, in the oracle's operation order: Delaunay meshes contain sliver
// elements whose Jacobian inversion amplifies any difference in rounding, and
// the FP64 assembly is to reproduce the CPU restatement to ~1e-12 on them too.
#pragma clang fp contract(off)

namespace svfsi {



// GNN for an affine tet4 (svFSI/NN.f:1469-1565): Nxi of TET4 are constant
// (NN.f:663-671): dN1 = e1, dN2 = e2, dN3 = e3, dN4 = -(1,1,1).
__device__ __forceinline__ void gnn_tet4(const double xl[4][3], double Nx[4][3], double& Jac,
                                         double ks[3][3])
{
    double A[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        A[i][0] = xl[0][i] - xl[3][i];
        A[i][1] = xl[1][i] - xl[3][i];
        A[i][2] = xl[2][i] - xl[3][i];
    }
    // the reference sums x_a * Nxi_a over a; for tet4 that is x_j - x_4
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
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) ks[i][j] = B[0][i] * B[0][j] + B[1][i] * B[1][j] + B[2][i] * B[2][j];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        Nx[0][j] = B[0][j];
        Nx[1][j] = B[1][j];
        Nx[2][j] = B[2][j];
        Nx[3][j] = -B[0][j] - B[1][j] - B[2][j];
    }
    Jac = J;
}

// ---------------------------------------------------------------------------
// LELAS3D (svFSI/LELAS.f:130-242) in CONSTRUCT_LELAS (LELAS.f:38-128), tet4.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_asm_lelas(int begin, int count, const int* __restrict__ ien,
                                                   const int* __restrict__ elemId,
                                                   const int* __restrict__ emap,
                                                   const double* __restrict__ x, int tDof,
                                                   const double* __restrict__ Ag,
                                                   const double* __restrict__ Dg,
                                                   const double* __restrict__ Eel, AsmParams P,
                                                   Tet4Rule rule, double* __restrict__ val,
                                                   double* __restrict__ R, double* __restrict__ ke,
                                                   double* __restrict__ re, const int* __restrict__ gslot,
                                                   const int* __restrict__ rslot)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * 16) return;
    const int p = begin + (t >> 4);
    const int ab = t & 15;
    const int a = ab >> 2, b = ab & 3;
    int nd[4];
    double xl[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        nd[q] = ien[(size_t)p * 4 + q];
#pragma unroll
        for (int i = 0; i < 3; ++i) xl[q][i] = x[(size_t)nd[q] * 3 + i];
    }
    double Nx[4][3], ks[3][3], Jac;
    gnn_tet4(xl, Nx, Jac, ks);
    const double E = Eel ? Eel[elemId[p]] : P.E;
    const double nu = P.nu, rho = P.rho;
    const double lambda = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu);
    const double mu = E * 0.5 / (1.0 + nu);
    const double lDm = lambda / mu;
    const double T1 = P.af * P.beta * P.dt * P.dt;
    const double amd = P.am / T1 * rho;
    double k[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double NxdNx = Nx[a][0] * Nx[b][0] + Nx[a][1] * Nx[b][1] + Nx[a][2] * Nx[b][2];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const double w = rule.w[g] * Jac;
        const double wl = w * T1 * mu;
        const double Tm = amd * rule.N[g][a] * rule.N[g][b] / mu + NxdNx;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double v = (i == j) ? Tm + (1.0 + lDm) * Nx[a][i] * Nx[b][i]
                                          : lDm * Nx[a][i] * Nx[b][j] + Nx[a][j] * Nx[b][i];
                k[i * 3 + j] += wl * v;
            }
    }
    if (ke) {
        double* kb = ke + (size_t)gslot[(size_t)p * 16 + ab] * 9;
#pragma unroll
        for (int q = 0; q < 9; ++q) kb[q] = k[q];
    } else {
        double* vb = val + (size_t)emap[(size_t)p * 16 + ab] * 9;
#pragma unroll
        for (int q = 0; q < 9; ++q) vb[q] += k[q];
    }

    if (b != 0) return;
    // residual row a (LELAS.f:161-205 and 208-215)
    double ed[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double* D = Dg + (size_t)nd[q] * tDof;
        ed[0] += Nx[q][0] * D[0];
        ed[1] += Nx[q][1] * D[1];
        ed[2] += Nx[q][2] * D[2];
        ed[3] += Nx[q][1] * D[0] + Nx[q][0] * D[1];
        ed[4] += Nx[q][2] * D[1] + Nx[q][1] * D[2];
        ed[5] += Nx[q][0] * D[2] + Nx[q][2] * D[0];
    }
    const double divD = lambda * (ed[0] + ed[1] + ed[2]);
    double S[6];
    S[0] = divD + 2.0 * mu * ed[0];
    S[1] = divD + 2.0 * mu * ed[1];
    S[2] = divD + 2.0 * mu * ed[2];
    S[3] = mu * ed[3];
    S[4] = mu * ed[4];
    S[5] = mu * ed[5];
    double r[3] = {0, 0, 0};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const double w = rule.w[g] * Jac;
        double ud[3] = {-P.fb[0], -P.fb[1], -P.fb[2]};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double* A = Ag + (size_t)nd[q] * tDof;
#pragma unroll
            for (int i = 0; i < 3; ++i) ud[i] += rule.N[g][q] * A[i];
        }
        const double Na = rule.N[g][a];
        r[0] += w * (rho * Na * ud[0] + Nx[a][0] * S[0] + Nx[a][1] * S[3] + Nx[a][2] * S[5]);
        r[1] += w * (rho * Na * ud[1] + Nx[a][0] * S[3] + Nx[a][1] * S[1] + Nx[a][2] * S[4]);
        r[2] += w * (rho * Na * ud[2] + Nx[a][0] * S[5] + Nx[a][1] * S[4] + Nx[a][2] * S[2]);
    }
    if (re) {
        double* rr = re + (size_t)rslot[(size_t)p * 4 + a] * 3;
        rr[0] = r[0];
        rr[1] = r[1];
        rr[2] = r[2];
        return;
    }
    double* Rr = R + (size_t)ien[(size_t)p * 4 + a] * 3;
    Rr[0] += r[0];
    Rr[1] += r[1];
    Rr[2] += r[2];
}

// ---------------------------------------------------------------------------
// FLUID3D_M + FLUID3D_C (svFSI/FLUID.f:192-560, 813-1084) in CONSTRUCT_FLUID
// (FLUID.f:40-190), tet4 VMS (nFs = 1), Newtonian viscosity: GNNxx of an
// affine tet is zero, so uxx = d2u2 = 0, and mu_x = mu_g = 0 (GETVISCOSITY
// viscType_Const, FLUID.f:1700-1703).  The momentum blocks (1..12) and the
// continuity blocks (13..16) are disjoint, so one Gauss loop builds both in
// the same per-entry order as the reference's two loops.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool iszero_d(double a)
{
    const double eps = 2.220446049250313e-16;
    const double aa = fabs(a);
    const double nrm = aa > eps ? aa : eps;
    return (aa / nrm) < 10.0 * eps;
}

// Forward-mode dual number: value and derivative with respect to one
// parameter.  The fluid element routine below is written once over a
// scalar type T; T = double assembles, T = Dual with mu = (mu_e, 1) gives the
// exact derivative of the element tangent and residual with respect to the
// element viscosity.
struct Dual {
    double v, d;
};
__device__ __forceinline__ Dual mkd(double v) { return {v, 0.0}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b)
{
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator/(double a, Dual b)
{
    const double q = a / b.v;
    return {q, -q * b.d / b.v};
}
__device__ __forceinline__ Dual dsqrt(Dual a)
{
    const double r = sqrt(a.v);
    return {r, 0.5 * a.d / r};
}
__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ double val_of(double a) { return a; }
__device__ __forceinline__ double val_of(Dual a) { return a.v; }
template <typename T> __device__ __forceinline__ T scal(double v);
template <> __device__ __forceinline__ double scal<double>(double v) { return v; }
template <> __device__ __forceinline__ Dual scal<Dual>(double v) { return mkd(v); }

// Element data of one tet4 shared by the 16 (a,b) threads of an element.
struct FluidElem {
    int nd[4];
    double Nx[4][3], ks[3][3], Jac;
    double yl[4][4], al[4][3];
    double ux[3][3], px[3], divU;
};

__device__ __forceinline__ void fluid_elem(int p, const int* __restrict__ ien, const double* __restrict__ x,
                                           int tDof, const double* __restrict__ Ag,
                                           const double* __restrict__ Yg, FluidElem& E)
{
    double xl[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        E.nd[q] = ien[(size_t)p * 4 + q];
#pragma unroll
        for (int i = 0; i < 3; ++i) xl[q][i] = x[(size_t)E.nd[q] * 3 + i];
    }
    gnn_tet4(xl, E.Nx, E.Jac, E.ks);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int c = 0; c < 4; ++c) E.yl[q][c] = Yg[(size_t)E.nd[q] * tDof + c];
#pragma unroll
        for (int c = 0; c < 3; ++c) E.al[q][c] = Ag[(size_t)E.nd[q] * tDof + c];
    }
    // element-constant kinematics: ux(i,c) = sum Nx(i,q) y(c,q), px, divU
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int c = 0; c < 3; ++c) E.ux[i][c] = 0.0;
        E.px[i] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int c = 0; c < 3; ++c) E.ux[i][c] += E.Nx[q][i] * E.yl[q][c];
        }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int i = 0; i < 3; ++i) E.px[i] += E.Nx[q][i] * E.yl[q][3];
    E.divU = E.ux[0][0] + E.ux[1][1] + E.ux[2][2];
}

// Gauss-point quantities of FLUID3D_M/C shared by all (a,b) blocks of a
// tet4 element: u, u_dot (with body force), p, the VMS parameters tau_M,
// tau_C, tau_B and the fine-scale velocity u' (svFSI/FLUID.f:220-412).
template <typename T> struct FluidGP {
    double u[3], ud[3], pp;
    T tauM, tauC, tauB, up[3];
};

template <typename T>
__device__ __forceinline__ void fluid_gp(const FluidElem& E, T mu, const AsmParams& P, const Tet4Rule& rule, int g,
                                         FluidGP<T>& G)
{
    const double rho = P.rho;
    const double ctM = 1.0, ctC = 36.0;
    double kS0 = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) kS0 += E.ks[i][j] * E.ks[i][j];
    const T kS = ctC * kS0 * (mu / rho) * (mu / rho);
    const double kT = 4.0 * (ctM / P.dt) * (ctM / P.dt);
    const double trK = E.ks[0][0] + E.ks[1][1] + E.ks[2][2];
    double ud[3] = {-P.fb[0], -P.fb[1], -P.fb[2]};
    double u[3] = {0, 0, 0};
    double pp = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double Nq = rule.N[g][q];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ud[c] += Nq * E.al[q][c];
            u[c] += Nq * E.yl[q][c];
        }
        pp += Nq * E.yl[q][3];
    }
    double kU = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) kU += u[i] * u[j] * E.ks[i][j];
    const T tauM = 1.0 / (rho * dsqrt(kT + kU + kS));
    T up[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double rV = ud[c] + u[0] * E.ux[0][c] + u[1] * E.ux[1][c] + u[2] * E.ux[2][c];
        up[c] = -tauM * (rho * rV + E.px[c] - mu * 0.0);
    }
    const T tauC = 1.0 / (tauM * trK);
    T tauB = scal<T>(0.0);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) tauB = tauB + up[i] * up[j] * E.ks[i][j];
    if (iszero_d(val_of(tauB))) tauB = scal<T>(2.220446049250313e-16);
    tauB = rho / dsqrt(tauB);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        G.u[c] = u[c];
        G.ud[c] = ud[c];
        G.up[c] = up[c];
    }
    G.pp = pp;
    G.tauM = tauM;
    G.tauC = tauC;
    G.tauB = tauB;
}

// Viscosity of a tet4 element (GETVISCOSITY, svFSI/FLUID.f:1690-1739, at
// the shear rate gamma = sqrt(es:es / 2) of FLUID.f:359-373): the velocity
// gradient of a linear tet is constant, so mu and mu_g = (d mu / d gamma) /
// gamma are element constants.  mu_i is the element's (or the equation's)
// viscosity; the constant model returns it with mu_g = 0.  Casson raises
// gamma to its threshold lam below it, as the reference does.
__device__ __forceinline__ void fluid_visc(const FluidElem& E, const AsmParams& P, double mu_i, double& mu,
                                           double& mu_g)
{
    mu = mu_i;
    mu_g = 0.0;
    if (P.viscType != 697 && P.viscType != 696) return;
    double gam = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double es = E.ux[i][j] + E.ux[j][i];
            gam += es * es;
        }
    gam = sqrt(0.5 * gam);
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
    mu_g = iszero_d(gam) ? 0.0 : mux / gam;
}

// FLUID3D_M + FLUID3D_C (svFSI/FLUID.f:192-560, 813-1084), the Gauss point
// g contribution to the (a,b) block of one tet4 element, Newtonian
// viscosity mu, second derivatives zero: k (4x4 block, row-major) and, when
// want_r, the residual of node a.
// Rows a and b of the element arrays, selected without dynamic indexing
// (which would place the element record in scratch memory).
struct FluidAB {
    double Nxa[3], Nxb[3];
    double Na[4], Nb[4]; // N_a, N_b at the four Gauss points
    double esNxa[3], esNxb[3]; // esNx(c,a) = sum_j es(j,c) Nx(j,a) (FLUID.f:321-328)
};

__device__ __forceinline__ void fluid_ab(const FluidElem& E, const Tet4Rule& rule, int a, int b, FluidAB& S)
{
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        S.Nxa[i] = 0.0;
        S.Nxb[i] = 0.0;
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        S.Na[g] = 0.0;
        S.Nb[g] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (q == a) S.Nxa[i] = E.Nx[q][i];
            if (q == b) S.Nxb[i] = E.Nx[q][i];
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            if (q == a) S.Na[g] = rule.N[g][q];
            if (q == b) S.Nb[g] = rule.N[g][q];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double e0 = E.ux[0][c] + E.ux[c][0], e1 = E.ux[1][c] + E.ux[c][1], e2 = E.ux[2][c] + E.ux[c][2];
        S.esNxa[c] = e0 * S.Nxa[0] + e1 * S.Nxa[1] + e2 * S.Nxa[2];
        S.esNxb[c] = e0 * S.Nxb[0] + e1 * S.Nxb[1] + e2 * S.Nxb[2];
    }
}

template <typename T>
__device__ __forceinline__ void fluid_pair_gp(const FluidElem& E, T mu, double mug, const AsmParams& P,
                                              const Tet4Rule& rule, int g, const FluidGP<T>& G, const FluidAB& S,
                                              bool want_r, T k[16], T r[4])
{
    const double* Nxa = S.Nxa;
    const double* Nxb = S.Nxb;
    const double rho = P.rho;
    const double T1c = P.af * P.gam * P.dt;
    const double amd = P.am / T1c;
    const double NxNx = Nxa[0] * Nxb[0] + Nxa[1] * Nxb[1] + Nxa[2] * Nxb[2];
    const double w = rule.w[g] * E.Jac;
    const double wl = w * T1c;
    const double wr = w * rho;
    const double* u = G.u;
    const double* ud = G.ud;
    const T tauM = G.tauM, tauC = G.tauC, tauB = G.tauB;
    const T* up = G.up;
    const T ua[3] = {u[0] + up[0], u[1] + up[1], u[2] + up[2]};

    const double Na = S.Na[g], Nb = S.Nb[g];
    const double uNxa = u[0] * Nxa[0] + u[1] * Nxa[1] + u[2] * Nxa[2];
    const double uNxb = u[0] * Nxb[0] + u[1] * Nxb[1] + u[2] * Nxb[2];
    const T upNxa = up[0] * Nxa[0] + up[1] * Nxa[1] + up[2] * Nxa[2];
    const T upNxb = up[0] * Nxb[0] + up[1] * Nxb[1] + up[2] * Nxb[2];
    const T uaNxa = uNxa + upNxa;
    const double T1b = -rho * uNxb;   // updu(i,c,b) = delta_ic T1b

    // momentum u-u blocks (FLUID.f:494-545)
    const T T1 = mu * NxNx + rho * amd * Nb * (Na + rho * tauM * uaNxa) + rho * Na * (uNxb + upNxb)
               + tauB * upNxa * upNxb;
#pragma unroll
    for (int rr = 0; rr < 3; ++rr)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double updu = (rr == c) ? T1b : 0.0;
            const T T2 = mu * Nxa[c] * Nxb[rr] + tauC * Nxa[rr] * Nxb[c] + S.esNxa[rr] * mug * S.esNxb[c]
                       - rho * tauM * uaNxa * updu;
            k[rr * 4 + c] = k[rr * 4 + c] + wl * ((rr == c) ? T2 + T1 : T2);
        }
    // momentum u-p blocks (FLUID.f:548-557)
    {
        const T T1p = rho * tauM * uaNxa;
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) k[rr * 4 + 3] = k[rr * 4 + 3] - wl * (Nxa[rr] * Nb - Nxb[rr] * T1p);
    }
    // continuity p-u and p-p blocks (FLUID.f:1052-1081)
    {
        const double T1q = rho * amd * Nb;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double T2 = Nxa[c] * (T1b - T1q);
            k[12 + c] = k[12 + c] + wl * (Na * Nxb[c] - tauM * T2);
        }
        k[15] = k[15] + wl * tauM * NxNx;
    }
    if (want_r) {
        // momentum residual (FLUID.f:429-465) and continuity (1046-1049)
        const T pa = G.pp - tauC * E.divU;
        T rVb[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            rVb[c] = tauB * (up[0] * E.ux[0][c] + up[1] * E.ux[1][c] + up[2] * E.ux[2][c]);
        T rM[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double es = E.ux[i][c] + E.ux[c][i];
                rM[i][c] = mu * es - rho * up[c] * ua[i] + rVb[c] * up[i] - (i == c ? pa : scal<T>(0.0));
            }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T rV = ud[c] + ua[0] * E.ux[0][c] + ua[1] * E.ux[1][c] + ua[2] * E.ux[2][c];
            r[c] = r[c] + (wr * Na * rV
                           + w * (Nxa[0] * rM[0][c] + Nxa[1] * rM[1][c] + Nxa[2] * rM[2][c]));
        }
        r[3] = r[3] + w * (Na * E.divU - upNxa);
    }
}

// The whole (a,b) block: the four Gauss points in order.
template <typename T>
__device__ __forceinline__ void fluid_pair(const FluidElem& E, T mu, const AsmParams& P, const Tet4Rule& rule,
                                           int a, int b, bool want_r, T k[16], T r[4])
{
    FluidAB S;
    fluid_ab(E, rule, a, b, S);
#pragma unroll
    for (int q = 0; q < 16; ++q) k[q] = scal<T>(0.0);
#pragma unroll
    for (int q = 0; q < 4; ++q) r[q] = scal<T>(0.0);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        FluidGP<T> G;
        fluid_gp(E, mu, P, rule, g, G);
        fluid_pair_gp(E, mu, 0.0, P, rule, g, G, S, want_r, k, r);
    }
}

__global__ __launch_bounds__(256) void k_asm_fluid(int begin, int count, const int* __restrict__ ien,
                                                   const int* __restrict__ elemId,
                                                   const int* __restrict__ emap,
                                                   const double* __restrict__ x, int tDof,
                                                   const double* __restrict__ Ag,
                                                   const double* __restrict__ Yg,
                                                   const double* __restrict__ muEl, AsmParams P,
                                                   Tet4Rule rule, double* __restrict__ val,
                                                   double* __restrict__ R, double* __restrict__ ke,
                                                   double* __restrict__ re, const int* __restrict__ gslot,
                                                   const int* __restrict__ rslot)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * 16) return;
    const int p = begin + (t >> 4);
    const int ab = t & 15;
    const int a = ab >> 2, b = ab & 3;
    FluidElem E;
    fluid_elem(p, ien, x, tDof, Ag, Yg, E);
    double mu, mug;
    fluid_visc(E, P, muEl ? muEl[elemId[p]] : P.mu, mu, mug);
    double k[16], r[4];
    // the 16 lanes of an element (consecutive lanes of one wave; count * 16
    // threads, so a group is active or inactive as a whole) evaluate the
    // Gauss-point quantities once per point, lane (ab & 3) for point ab & 3,
    // and fetch each point's values from its lane
    FluidGP<double> mine;
    {
        // the point this lane evaluates, selected without dynamic indexing
        FluidGP<double> G;
#pragma unroll
        for (int g = 0; g < 4; ++g)
            if (g == (ab & 3)) fluid_gp<double>(E, mu, P, rule, g, mine);
        (void)G;
    }
    FluidAB S;
    fluid_ab(E, rule, a, b, S);
    const int base = (threadIdx.x & (kWave - 1)) & ~15;
#pragma unroll
    for (int q = 0; q < 16; ++q) k[q] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) r[q] = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        FluidGP<double> G;
        const int src = base + g;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            G.u[c] = __shfl(mine.u[c], src, kWave);
            G.ud[c] = __shfl(mine.ud[c], src, kWave);
            G.up[c] = __shfl(mine.up[c], src, kWave);
        }
        G.pp = __shfl(mine.pp, src, kWave);
        G.tauM = __shfl(mine.tauM, src, kWave);
        G.tauC = __shfl(mine.tauC, src, kWave);
        G.tauB = __shfl(mine.tauB, src, kWave);
        fluid_pair_gp<double>(E, mu, mug, P, rule, g, G, S, b == 0, k, r);
    }
    if (ke) {
        double* kb = ke + (size_t)gslot[(size_t)p * 16 + ab] * 16;
#pragma unroll
        for (int q = 0; q < 16; ++q) kb[q] = k[q];
        if (b == 0) {
            double* rr = re + (size_t)rslot[(size_t)p * 4 + a] * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) rr[c] = r[c];
        }
        return;
    }
    double* vb = val + (size_t)emap[(size_t)p * 16 + ab] * 16;
#pragma unroll
    for (int q = 0; q < 16; ++q) vb[q] += k[q];
    if (b == 0) {
        // the node index reloaded from ien: indexing the per-thread E.nd with
        // the lane-dependent a would place the whole element record in
        // scratch memory
        double* Rr = R + (size_t)ien[(size_t)p * 4 + a] * 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) Rr[c] += r[c];
    }
}

// Viscosity sensitivity of the fluid element (no reference counterpart):
// g_e = lam_e^T (dR_e/dmu_e - dK_e/dmu_e du_e), the element term of
// dL/dmu_e when K du = R and K^T lam = dL/du.  Sixteen lanes per element
// (one per (a,b) block) evaluate the dual-number element routine; the
// element sum is a fixed-order shuffle reduction over the 16 lanes.
__global__ __launch_bounds__(256) void k_fluid_dmu(int nEl, const int* __restrict__ ien,
                                                   const int* __restrict__ elemId, const double* __restrict__ x,
                                                   int tDof, const double* __restrict__ Ag,
                                                   const double* __restrict__ Yg, const double* __restrict__ muEl,
                                                   AsmParams P, Tet4Rule rule, const double* __restrict__ lam,
                                                   const double* __restrict__ du, double* __restrict__ g)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nEl * 16) return;
    const int p = t >> 4;
    const int ab = t & 15;
    const int a = ab >> 2, b = ab & 3;
    FluidElem E;
    fluid_elem(p, ien, x, tDof, Ag, Yg, E);
    const Dual mu = {muEl ? muEl[elemId[p]] : P.mu, 1.0};
    Dual k[16], r[4];
    fluid_pair<Dual>(E, mu, P, rule, a, b, b == 0, k, r);
    const double* la = lam + (size_t)ien[(size_t)p * 4 + a] * 4;
    const double* ub = du + (size_t)ien[(size_t)p * 4 + b] * 4;
    double s = 0.0;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
        double kd = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) kd += k[rr * 4 + c].d * ub[c];
        s += la[rr] * ((b == 0 ? r[rr].d : 0.0) - kd);
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
    if (ab == 0) g[elemId[p]] = s;
}

// ---------------------------------------------------------------------------
// Face terms on tri3 faces (GNNB, svFSI/NN.f:1856-1996).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void gnnb_tri3(const double* x, int n0, int n1, int n2, int inner,
                                          double n[3])
{
    double t1[3], t2[3], v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        t1[i] = x[(size_t)n0 * 3 + i] - x[(size_t)n2 * 3 + i];
        t2[i] = x[(size_t)n1 * 3 + i] - x[(size_t)n2 * 3 + i];
        v[i] = x[(size_t)n0 * 3 + i] - x[(size_t)inner * 3 + i];
    }
    n[0] = t1[1] * t2[2] - t1[2] * t2[1];
    n[1] = t1[2] * t2[0] - t1[0] * t2[2];
    n[2] = t1[0] * t2[1] - t1[1] * t2[0];
    if (n[0] * v[0] + n[1] * v[1] + n[2] * v[2] < 0.0) {
        n[0] = -n[0];
        n[1] = -n[1];
        n[2] = -n[2];
    }
}

// SETBCTRACL (svFSI/SETBC.f:319-416): lR(:,a) -= w N_a h ; one thread per
// (face, a).
__global__ void k_face_traction(int begin, int count, const int* __restrict__ order,
                                const int* __restrict__ fien, const int* __restrict__ inner,
                                const double* __restrict__ x, int dof, double h0, double h1,
                                double h2, Tri3Rule rule, double* __restrict__ R)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * 3) return;
    const int f = order[begin + t / 3];
    const int a = t % 3;
    const int n0 = fien[f * 3], n1 = fien[f * 3 + 1], n2 = fien[f * 3 + 2];
    double n[3];
    gnnb_tri3(x, n0, n1, n2, inner[f], n);
    const double Jac = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double h[3] = {h0, h1, h2};
    double r[3] = {0, 0, 0};
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        const double w = rule.w[g] * Jac;
#pragma unroll
        for (int i = 0; i < 3; ++i) r[i] -= w * rule.N[g][a] * h[i];
    }
    const int node = fien[f * 3 + a];
#pragma unroll
    for (int i = 0; i < 3; ++i) R[(size_t)node * dof + i] += r[i];
}

// BFLUID with h = 0 (svFSI/FLUID.f:1280-1336); one thread per (face, a, b).
__global__ void k_face_backflow(int begin, int count, const int* __restrict__ order,
                                const int* __restrict__ fien, const int* __restrict__ inner,
                                const int* __restrict__ fmap, const double* __restrict__ x,
                                int tDof, const double* __restrict__ Yg, AsmParams P,
                                Tri3Rule rule, double* __restrict__ val, double* __restrict__ R)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * 9) return;
    const int f = order[begin + t / 9];
    const int ab = t % 9;
    const int a = ab / 3, b = ab % 3;
    const int nd[3] = {fien[f * 3], fien[f * 3 + 1], fien[f * 3 + 2]};
    double n[3];
    gnnb_tri3(x, nd[0], nd[1], nd[2], inner[f], n);
    const double Jac = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double nV[3] = {n[0] / Jac, n[1] / Jac, n[2] / Jac};
    double kd = 0.0;
    double r[3] = {0, 0, 0};
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        const double w = rule.w[g] * Jac;
        double y[3] = {0, 0, 0};
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int i = 0; i < 3; ++i) y[i] += rule.N[g][q] * Yg[(size_t)nd[q] * tDof + i];
        const double wl = w * P.af * P.gam * P.dt;
        double udn = y[0] * nV[0] + y[1] * nV[1] + y[2] * nV[2];
        udn = 0.5 * P.bstab * P.rho * (udn - fabs(udn));
        kd -= wl * rule.N[g][a] * rule.N[g][b] * udn;
        if (b == 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i) r[i] -= w * rule.N[g][a] * (udn * y[i]);
        }
    }
    double* vb = val + (size_t)fmap[f * 9 + ab] * 16;
    vb[0] += kd;
    vb[5] += kd;
    vb[10] += kd;
    if (b == 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) R[(size_t)nd[a] * 4 + i] += r[i];
    }
}

// ---------------------------------------------------------------------------
// Adjoint: g_e = -lam_e^T (dK_e/dE_e) u_e for the LELAS tet4 tangent.  The
// stiffness part of LELAS3D is proportional to E (lambda, mu ~ E) and the
// mass part does not depend on E, so dK_e/dE_e is the stiffness tangent at
// unit modulus.  One thread per element; element matrices are recomputed,
// never stored.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lelas_dkde(int nEl, const int* __restrict__ ien,
                                                    const int* __restrict__ elemId,
                                                    const double* __restrict__ x, AsmParams P,
                                                    Tet4Rule rule, const double* __restrict__ lam,
                                                    const double* __restrict__ u,
                                                    double* __restrict__ g)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nEl) return;
    int nd[4];
    double xl[4][3], la[4][3], ub[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        nd[q] = ien[(size_t)p * 4 + q];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            xl[q][i] = x[(size_t)nd[q] * 3 + i];
            la[q][i] = lam[(size_t)nd[q] * 3 + i];
            ub[q][i] = u[(size_t)nd[q] * 3 + i];
        }
    }
    double Nx[4][3], ks[3][3], Jac;
    gnn_tet4(xl, Nx, Jac, ks);
    const double nu = P.nu;
    const double lambda = nu / (1.0 + nu) / (1.0 - 2.0 * nu);
    const double mu = 0.5 / (1.0 + nu);
    const double lDm = lambda / mu;
    const double T1 = P.af * P.beta * P.dt * P.dt;
    double wsum = 0.0;
#pragma unroll
    for (int gq = 0; gq < 4; ++gq) wsum += rule.w[gq] * Jac * T1 * mu;
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const double NxdNx = Nx[a][0] * Nx[b][0] + Nx[a][1] * Nx[b][1] + Nx[a][2] * Nx[b][2];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double v = (i == j) ? NxdNx + (1.0 + lDm) * Nx[a][i] * Nx[b][i]
                                              : lDm * Nx[a][i] * Nx[b][j] + Nx[a][j] * Nx[b][i];
                    s += la[a][i] * (wsum * v) * ub[b][j];
                }
        }
    g[elemId[p]] = -s;
}

// ---------------------------------------------------------------------------
// STRUCT3D + GETPK2CC nHK (svFSI/STRUCT.f:164-496, MATMODELS.f:41-150,
// 539-564, 1588-1647) in CONSTRUCT_dSOLID (STRUCT.f:39-162), TET10, solid
// viscosity 0, no damping, prestress or fibres.  One wavefront handles EPW
// elements of one colour.  Phase 1: a lane per (element, Gauss point)
// evaluates GNN, F, the neo-Hookean S and Voigt Dm, P = F S and the inertia
// term into an LDS record.  Phase 2: lanes stride over the (element, a, b)
// node pairs, accumulate the 3x3 block over the Gauss points as the
// reference does for that block, and add it to Val through the scatter map;
// the b = 0 lane also builds the residual row of node a.
// ---------------------------------------------------------------------------
// Gauss-point record layout (doubles).  The shape functions come last: the
// material-point records kept in HBM (SR_HBM doubles each) leave them out,
// since N depends only on the rule.
constexpr int SR_W = 0, SR_NX = 1, SR_S = 31, SR_DM = 40, SR_P = 76, SR_UD = 85, SR_F = 88, SR_HBM = 97,
              SR_N = 97, SR_SZ = 107;

__device__ __forceinline__ double det3d(const double A[3][3])
{
    return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1])
         - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
         + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}

// Material point (element e, Gauss point with rule record rg) of STRUCT3D:
// GNN, F, the neo-Hookean S and Voigt Dm (GETPK2CC), P = F S and the inertia
// term, written to the record rc (layout SR_*; N only when WITH_N).
template <bool WITH_N>
__device__ __forceinline__ void struct_point(int e, const int* __restrict__ ien, const double* __restrict__ x,
                                             int tDof, const double* __restrict__ Ag,
                                             const double* __restrict__ Dg, const AsmParams& P,
                                             const double* __restrict__ rg, double* rc)
{
    // GNN (NN.f:1469-1565), streaming the nodes to keep registers low
    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll 2
    for (int a = 0; a < 10; ++a) {
        const int na = ien[(size_t)e * 10 + a];
        double xa[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) xa[i] = x[(size_t)na * 3 + i];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) A[i][j] += xa[i] * rg[11 + a * 3 + j];
    }
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
    // Nx (to LDS), F and inertia (STRUCT.f:195-235)
    double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double ud[3] = {-P.rho * P.fb[0], -P.rho * P.fb[1], -P.rho * P.fb[2]};
#pragma unroll 2
    for (int a = 0; a < 10; ++a) {
        const int na = ien[(size_t)e * 10 + a];
        double Nxa[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Nxa[j] = rg[11 + a * 3] * B[0][j] + rg[11 + a * 3 + 1] * B[1][j] + rg[11 + a * 3 + 2] * B[2][j];
            rc[SR_NX + a * 3 + j] = Nxa[j];
        }
        const double Na = rg[1 + a];
        if constexpr (WITH_N) rc[SR_N + a] = Na;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double Dai = Dg[(size_t)na * tDof + i];
            ud[i] += Na * (P.rho * Ag[(size_t)na * tDof + i]);
#pragma unroll
            for (int j = 0; j < 3; ++j) F[i][j] += Nxa[j] * Dai;
        }
    }
    // GETPK2CC, neo-Hookean + GETSVOLP
    const double nd3 = 3.0;
    const double Jf = det3d(F);
    const double J2d = pow(Jf, -2.0 / nd3);
    double C[3][3], Ci[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
    {
        const double d = det3d(C);
        Ci[0][0] = (C[1][1] * C[2][2] - C[1][2] * C[2][1]) / d;
        Ci[0][1] = (C[0][2] * C[2][1] - C[0][1] * C[2][2]) / d;
        Ci[0][2] = (C[0][1] * C[1][2] - C[0][2] * C[1][1]) / d;
        Ci[1][0] = (C[1][2] * C[2][0] - C[1][0] * C[2][2]) / d;
        Ci[1][1] = (C[0][0] * C[2][2] - C[0][2] * C[2][0]) / d;
        Ci[1][2] = (C[0][2] * C[1][0] - C[0][0] * C[1][2]) / d;
        Ci[2][0] = (C[1][0] * C[2][1] - C[1][1] * C[2][0]) / d;
        Ci[2][1] = (C[0][1] * C[2][0] - C[0][0] * C[2][1]) / d;
        Ci[2][2] = (C[0][0] * C[1][1] - C[0][1] * C[1][0]) / d;
    }
    double p = 0.0, pl = 0.0;
    const double Kp = P.Kpen;
    if (!iszero_d(Kp)) {
        if (P.volType == 1) { p = Kp * (Jf - 1.0); pl = Kp * (2.0 * Jf - 1.0); }
        else if (P.volType == 3) { p = Kp * (1.0 - 1.0 / Jf); pl = Kp; }
        else { p = 0.5 * Kp * (Jf - 1.0 / Jf); pl = Kp * Jf; }
    }
    const double g1 = 2.0 * P.C10;
    double CSb = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) CSb += C[i][j] * (i == j ? g1 : 0.0);
    const double r1 = J2d * CSb / nd3;
    double Sd[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Sd[i][j] = J2d * (i == j ? g1 : 0.0) - r1 * Ci[i][j];
    const int vi[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {1, 2}, {2, 0}};
#pragma unroll
    for (int aa = 0; aa < 6; ++aa)
#pragma unroll
        for (int bb = 0; bb < 6; ++bb) {
            const int i = vi[aa][0], j = vi[aa][1], kk = vi[bb][0], l = vi[bb][1];
            const double dy1 = Ci[i][j] * Sd[kk][l] + Sd[i][j] * Ci[kk][l];
            const double sym = 0.5 * (Ci[i][kk] * Ci[j][l] + Ci[i][l] * Ci[j][kk]);
            const double dyc = Ci[i][j] * Ci[kk][l];
            rc[SR_DM + aa * 6 + bb] = (-2.0 / nd3) * dy1 + 2.0 * (r1 - p * Jf) * sym + (pl * Jf - 2.0 * r1 / nd3) * dyc;
        }
    double S[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            S[i][j] = Sd[i][j] + p * Jf * Ci[i][j];
            rc[SR_S + i * 3 + j] = S[i][j];
            rc[SR_F + i * 3 + j] = F[i][j];
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) rc[SR_P + i * 3 + j] = F[i][0] * S[0][j] + F[i][1] * S[1][j] + F[i][2] * S[2][j];
    rc[SR_W] = rg[0] * J;
#pragma unroll
    for (int i = 0; i < 3; ++i) rc[SR_UD + i] = ud[i];
}

template <int NG, int EPW>
__global__ __launch_bounds__(256) void k_asm_struct(int begin, int count, const int* __restrict__ ien,
                                                   const int* __restrict__ emap, const double* __restrict__ x,
                                                   int tDof, const double* __restrict__ Ag,
                                                   const double* __restrict__ Dg, AsmParams P,
                                                   const double* __restrict__ rule, double* __restrict__ val,
                                                   double* __restrict__ R)
{
    __shared__ double rec[EPW * NG][SR_SZ];
    __shared__ double dbm[EPW * NG][10][18]; // Dm B_b per (element, Gauss point, node b)
    const int lane = threadIdx.x;
    const int e0 = begin + blockIdx.x * EPW;
    const int eEnd = begin + count;
    if (lane < EPW * NG) {
        const int k = lane / NG, g = lane - (lane / NG) * NG;
        const int e = e0 + k;
        double* rc = rec[lane];
        if (e < eEnd) struct_point<true>(e, ien, x, tDof, Ag, Dg, P, rule + g * kTet10Rec, rc);
    }
    __syncthreads();
    // Dm B_b, shared by the ten blocks (a, b) of column b
    for (int idx = lane; idx < EPW * NG * 10; idx += 256) {
        const int kg = idx / 10, b = idx - (idx / 10) * 10;
        if (e0 + kg / NG >= eEnd) continue;
        const double* rc = rec[kg];
        const double* Nxb = rc + SR_NX + b * 3;
        const double* Fg = rc + SR_F;
        double Bb[6][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Bb[0][c] = Nxb[0] * Fg[c * 3 + 0];
            Bb[1][c] = Nxb[1] * Fg[c * 3 + 1];
            Bb[2][c] = Nxb[2] * Fg[c * 3 + 2];
            Bb[3][c] = Nxb[0] * Fg[c * 3 + 1] + Fg[c * 3 + 0] * Nxb[1];
            Bb[4][c] = Nxb[1] * Fg[c * 3 + 2] + Fg[c * 3 + 1] * Nxb[2];
            Bb[5][c] = Nxb[2] * Fg[c * 3 + 0] + Fg[c * 3 + 2] * Nxb[0];
        }
#pragma unroll
        for (int rr = 0; rr < 6; ++rr)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double sacc = 0.0;
#pragma unroll
                for (int q = 0; q < 6; ++q) sacc += rc[SR_DM + rr * 6 + q] * Bb[q][c];
                dbm[kg][b][rr * 3 + c] = sacc;
            }
    }
    __syncthreads();
    const double afu = P.af * P.beta * P.dt * P.dt;
    const double amd = P.am * P.rho;
    for (int pidx = lane; pidx < EPW * 100; pidx += 256) {
        const int k = pidx / 100, ab = pidx - (pidx / 100) * 100;
        const int a = ab / 10, b = ab - (ab / 10) * 10;
        const int e = e0 + k;
        if (e >= eEnd) continue;
        double kk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double r[3] = {0, 0, 0};
#pragma unroll 1
        for (int g = 0; g < NG; ++g) {
            const double* rc = rec[k * NG + g];
            const double w = rc[SR_W];
            const double* Nxa = rc + SR_NX + a * 3;
            const double* Nxb = rc + SR_NX + b * 3;
            const double* Sg = rc + SR_S;
            const double* Fg = rc + SR_F;
            double NxSNx = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i) NxSNx += Nxa[i] * Sg[i * 3 + j] * Nxb[j];
            const double T1 = amd * rc[SR_N + a] * rc[SR_N + b] + afu * NxSNx;
            double Ba[6][3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                Ba[0][c] = Nxa[0] * Fg[c * 3 + 0];
                Ba[1][c] = Nxa[1] * Fg[c * 3 + 1];
                Ba[2][c] = Nxa[2] * Fg[c * 3 + 2];
                Ba[3][c] = Nxa[0] * Fg[c * 3 + 1] + Fg[c * 3 + 0] * Nxa[1];
                Ba[4][c] = Nxa[1] * Fg[c * 3 + 2] + Fg[c * 3 + 1] * Nxa[2];
                Ba[5][c] = Nxa[2] * Fg[c * 3 + 0] + Fg[c * 3 + 2] * Nxa[0];
            }
            const double* DBm = dbm[k * NG + g][b];
#pragma unroll
            for (int rr = 0; rr < 3; ++rr)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double BmDBm = 0.0;
#pragma unroll
                    for (int q = 0; q < 6; ++q) BmDBm += Ba[q][rr] * DBm[q * 3 + c];
                    kk[rr * 3 + c] += w * ((rr == c ? T1 : 0.0) + afu * BmDBm);
                }
            if (b == 0) {
                const double Na = rc[SR_N + a];
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    r[i] += w * (Na * rc[SR_UD + i] + Nxa[0] * rc[SR_P + i * 3] + Nxa[1] * rc[SR_P + i * 3 + 1]
                                 + Nxa[2] * rc[SR_P + i * 3 + 2]);
            }
        }
        double* vb = val + (size_t)emap[(size_t)e * 100 + ab] * 9;
#pragma unroll
        for (int q = 0; q < 9; ++q) vb[q] += kk[q];
        if (b == 0) {
            double* Rr = R + (size_t)ien[(size_t)e * 10 + a] * 3;
#pragma unroll
            for (int i = 0; i < 3; ++i) Rr[i] += r[i];
        }
    }
}

// ---------------------------------------------------------------------------
// STRUCT3D with the B^T Dm B contraction on the matrix cores.
//
// k_struct_points evaluates every material point (element, Gauss point) of
// the mesh in one launch, a lane each, and stores the record (SR_HBM doubles,
// colour-sorted element order) in HBM.  k_asm_struct_mfma then assembles one
// colour: per element the material stiffness
//   K_mat[(a,i),(b,j)] = afu sum_g w_g (B_a,g^T Dm_g B_b,g)_ij
// is the product A^T C of two (6 nG) x 30 matrices, A[(g,q),(a,i)] =
// B_a,g[q][i] and C[(g,q),(b,j)] = afu w_g (Dm_g B_b,g)[q][j] (STRUCT.f:
// 295-319, 348-493), padded to 32 columns and a multiple of 4 rows.  Each
// of its four 16 x 16 output tiles is a chain of 16x16x4 MFMA steps (FP64,
// or FP32 with FP32 accumulation).  The geometric and inertia term
// delta_ij w_g (amd N_a N_b + afu Nx_a.S.Nx_b) and the residual stay on the
// FP64 vector units in the scatter pass.
// ---------------------------------------------------------------------------
template <int NG>
__global__ __launch_bounds__(256) void k_struct_points(int nEl, const int* __restrict__ ien,
                                                      const double* __restrict__ x, int tDof,
                                                      const double* __restrict__ Ag,
                                                      const double* __restrict__ Dg, AsmParams P,
                                                      const double* __restrict__ rule, double* __restrict__ rec)
{
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (size_t)nEl * NG) return;
    const int e = (int)(t / NG), g = (int)(t - (size_t)e * NG);
    struct_point<false>(e, ien, x, tDof, Ag, Dg, P, rule + g * kTet10Rec, rec + t * SR_HBM);
}

// One 16x16x4 MFMA step.  Operands: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16].  The output item r of lane l is D[row(l, r)][l % 16];
// the row order differs between the FP64 and FP32 instructions (measured
// on gfx950 with random operands).
template <typename MT> struct MfmaTile;
template <> struct MfmaTile<double> {
    typedef double V __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V step(double a, double b, V c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};
template <> struct MfmaTile<float> {
    typedef float V __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V step(float a, float b, V c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

template <int NG, typename MT, int EPW>
__global__ __launch_bounds__(256) void k_asm_struct_mfma(int begin, int count, const int* __restrict__ ien,
                                                        const int* __restrict__ emap, AsmParams P,
                                                        const double* __restrict__ rule,
                                                        const double* __restrict__ rec, double* __restrict__ val,
                                                        double* __restrict__ R, double* __restrict__ ke,
                                                        double* __restrict__ re, const int* __restrict__ gslot,
                                                        const int* __restrict__ rslot)
{
    constexpr int KP = (6 * NG + 3) / 4 * 4;   // contraction length, padded to the MFMA depth
    constexpr int OPN = 2 * EPW * KP * 32 > EPW * 1024 ? 2 * EPW * KP * 32 : EPW * 1024;
    constexpr int NTASK = EPW * 4;             // (element, output tile) pairs
    constexpr int TPW = (NTASK + 3) / 4;       // per wave
    typedef typename MfmaTile<MT>::V Acc;
    __shared__ double rs[EPW * NG * SR_HBM];
    __shared__ MT op[OPN];
    MT* Aop = op;                    // [EPW][KP][32]
    MT* Cop = op + EPW * KP * 32;    // [EPW][KP][32]
    MT* Kout = op;                   // [EPW][32][32], after the products
    const int tid = threadIdx.x;
    const int e0 = begin + blockIdx.x * EPW;
    const int nE = min(EPW, begin + count - e0);
    {
        const double* src = rec + (size_t)e0 * NG * SR_HBM;
        for (int i = tid; i < nE * NG * SR_HBM; i += 256) rs[i] = src[i];
    }
    // zero padding: columns 30, 31 of every row, and rows 6 NG .. KP-1
    for (int i = tid; i < EPW * KP * 2; i += 256) {
        const int row = i >> 1, c = 30 + (i & 1);
        Aop[row * 32 + c] = (MT)0;
        Cop[row * 32 + c] = (MT)0;
    }
    if constexpr (KP > 6 * NG) {
        for (int i = tid; i < EPW * (KP - 6 * NG) * 32; i += 256) {
            const int e = i / ((KP - 6 * NG) * 32), r = i - e * ((KP - 6 * NG) * 32);
            Aop[((size_t)e * KP + 6 * NG) * 32 + r] = (MT)0;
            Cop[((size_t)e * KP + 6 * NG) * 32 + r] = (MT)0;
        }
    }
    __syncthreads();
    const double afu = P.af * P.beta * P.dt * P.dt;
    // operands: B_b (rows (g, q), columns (b, j)) and afu w Dm B_b
    for (int t = tid; t < nE * NG * 10; t += 256) {
        const int e = t / (NG * 10), r = t - e * (NG * 10);
        const int g = r / 10, b = r - g * 10;
        const double* rc = rs + (e * NG + g) * SR_HBM;
        const double* Nxb = rc + SR_NX + b * 3;
        const double* Fg = rc + SR_F;
        double Bb[6][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Bb[0][c] = Nxb[0] * Fg[c * 3 + 0];
            Bb[1][c] = Nxb[1] * Fg[c * 3 + 1];
            Bb[2][c] = Nxb[2] * Fg[c * 3 + 2];
            Bb[3][c] = Nxb[0] * Fg[c * 3 + 1] + Fg[c * 3 + 0] * Nxb[1];
            Bb[4][c] = Nxb[1] * Fg[c * 3 + 2] + Fg[c * 3 + 1] * Nxb[2];
            Bb[5][c] = Nxb[2] * Fg[c * 3 + 0] + Fg[c * 3 + 2] * Nxb[0];
        }
        const double sc = afu * rc[SR_W];
        MT* Ar = Aop + ((size_t)e * KP + g * 6) * 32 + b * 3;
        MT* Cr = Cop + ((size_t)e * KP + g * 6) * 32 + b * 3;
#pragma unroll
        for (int q = 0; q < 6; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double d = 0.0;
#pragma unroll
                for (int m = 0; m < 6; ++m) d += rc[SR_DM + q * 6 + m] * Bb[m][c];
                Ar[q * 32 + c] = (MT)Bb[q][c];
                Cr[q * 32 + c] = (MT)(sc * d);
            }
    }
    __syncthreads();
    // products: wave w takes the (element, tile) tasks w, w + 4, ...
    const int wave = tid >> 6, l = tid & 63;
    Acc acc[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        acc[u] = Acc{0, 0, 0, 0};
        const int task = wave + 4 * u;
        const int e = task >> 2;
        if (task < NTASK && e < nE) { // uniform across the wave
            const int ti = (task >> 1) & 1, tj = task & 1;
            const MT* Ab = Aop + (size_t)e * KP * 32 + (l >> 4) * 32 + 16 * ti + (l & 15);
            const MT* Cb = Cop + (size_t)e * KP * 32 + (l >> 4) * 32 + 16 * tj + (l & 15);
#pragma unroll 6
            for (int s = 0; s < KP / 4; ++s) acc[u] = MfmaTile<MT>::step(Ab[s * 128], Cb[s * 128], acc[u]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        const int task = wave + 4 * u;
        const int e = task >> 2;
        if (task < NTASK && e < nE) {
            const int ti = (task >> 1) & 1, tj = task & 1;
            MT* Kt = Kout + e * 1024 + 16 * ti * 32 + 16 * tj + (l & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) Kt[MfmaTile<MT>::row(l, r) * 32] = acc[u][r];
        }
    }
    __syncthreads();
    // scatter: geometric + inertia term and the residual in FP64, then Val
    const double amd = P.am * P.rho;
    for (int t = tid; t < nE * 100; t += 256) {
        const int e = t / 100, ab = t - (t / 100) * 100;
        const int a = ab / 10, b = ab - (ab / 10) * 10;
        double T1 = 0.0;
        double r[3] = {0, 0, 0};
#pragma unroll 1
        for (int g = 0; g < NG; ++g) {
            const double* rc = rs + (e * NG + g) * SR_HBM;
            const double* rg = rule + g * kTet10Rec;
            const double w = rc[SR_W];
            const double* Nxa = rc + SR_NX + a * 3;
            const double* Nxb = rc + SR_NX + b * 3;
            const double* Sg = rc + SR_S;
            double NxSNx = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i) NxSNx += Nxa[i] * Sg[i * 3 + j] * Nxb[j];
            const double Na = rg[1 + a];
            T1 += w * (amd * Na * rg[1 + b] + afu * NxSNx);
            if (b == 0) {
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    r[i] += w * (Na * rc[SR_UD + i] + Nxa[0] * rc[SR_P + i * 3] + Nxa[1] * rc[SR_P + i * 3 + 1]
                                 + Nxa[2] * rc[SR_P + i * 3 + 2]);
            }
        }
        const MT* Kb = Kout + e * 1024 + (3 * a) * 32 + 3 * b;
        if (ke) {
            double* kb = ke + (size_t)gslot[(size_t)(e0 + e) * 100 + ab] * 9;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) kb[i * 3 + j] = (double)Kb[i * 32 + j] + (i == j ? T1 : 0.0);
            if (b == 0) {
                double* rr = re + (size_t)rslot[(size_t)(e0 + e) * 10 + a] * 3;
#pragma unroll
                for (int i = 0; i < 3; ++i) rr[i] = r[i];
            }
            continue;
        }
        double* vb = val + (size_t)emap[(size_t)(e0 + e) * 100 + ab] * 9;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) vb[i * 3 + j] += (double)Kb[i * 32 + j] + (i == j ? T1 : 0.0);
        if (b == 0) {
            double* Rr = R + (size_t)ien[(size_t)(e0 + e) * 10 + a] * 3;
#pragma unroll
            for (int i = 0; i < 3; ++i) Rr[i] += r[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Gather assembly: Val(:, j) += the sum of the element blocks that land on
// nnz block j, R(:, i) += the sum of node i's element residual rows.  The
// element kernels stored those contributions contiguously, in element-id
// order (DOASSEM's element loop, LHSA.f:266-298), so each sum is a streaming
// read, deterministic, and associates like the reference's.  DD consecutive
// lanes share one block.
// ---------------------------------------------------------------------------
template <int DD>
__global__ __launch_bounds__(256) void k_gather(size_t nItems, const int* __restrict__ ptr,
                                               const double* __restrict__ src, double* __restrict__ dst)
{
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t j = t / DD;
    if (j >= nItems) return;
    const int q = (int)(t - j * DD);
    const int k1 = ptr[j + 1];
    double sum = 0.0;
    int k = ptr[j];
    // four loads in flight, added in element order (the sums associate as
    // DOASSEM's element loop does)
    for (; k + 4 <= k1; k += 4) {
        const double a0 = src[(size_t)k * DD + q], a1 = src[(size_t)(k + 1) * DD + q];
        const double a2 = src[(size_t)(k + 2) * DD + q], a3 = src[(size_t)(k + 3) * DD + q];
        sum += a0;
        sum += a1;
        sum += a2;
        sum += a3;
    }
    for (; k < k1; ++k) sum += src[(size_t)k * DD + q];
    dst[j * DD + q] += sum;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// Gather assembly: one launch over all elements stores the element blocks
// at their gather positions, then k_gather sums them into Val and R.  The
// default for tet4, whose 16- and 9-double blocks; TET10 (STRUCT) keeps the
// coloured scatter, which measured faster there (5.0 vs 5.8 ms on the 40^3
// tet10 cube: its 72-byte blocks straddle cache lines, so the scattered
// stores of the element blocks cost more than the coloured read-modify-
// write).  SVFSI_ASM_GATHER = 0 / 1 overrides.
static bool gather_enabled(int eNoN)
{
    const char* v = getenv("SVFSI_ASM_GATHER");
    if (v && (v[0] == '0' || v[0] == '1')) return v[0] == '1';
    return eNoN == 4;
}

int launch_gather(Mesh& M, int dof, double* val, double* R, hipStream_t s)
{
    const size_t nb = (size_t)M.lhs->nnz, nr = (size_t)M.nNo;
    const double* ke = M.ke.as<double>();
    const double* re = M.re.as<double>();
    auto grid = [](size_t items, int dd) { return dim3((unsigned)((items * dd + 255) / 256)); };
    if (dof == 3) {
        hipLaunchKernelGGL(k_gather<9>, grid(nb, 9), dim3(256), 0, s, nb, M.gPtr.as<int>(), ke, val);
        hipLaunchKernelGGL(k_gather<3>, grid(nr, 3), dim3(256), 0, s, nr, M.rPtr.as<int>(), re, R);
    } else {
        hipLaunchKernelGGL(k_gather<16>, grid(nb, 16), dim3(256), 0, s, nb, M.gPtr.as<int>(), ke, val);
        hipLaunchKernelGGL(k_gather<4>, grid(nr, 4), dim3(256), 0, s, nr, M.rPtr.as<int>(), re, R);
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

int launch_asm_tet4(Mesh& M, int phys, const AsmParams& P, const Tet4Rule& rule, int tDof, const double* Ag,
                    const double* Yg, const double* Dg, const double* prop, double* val, double* R, hipStream_t s)
{
    const int dof = phys == SVFSI_PHYS_LELAS ? 3 : 4;
    if (gather_enabled(M.eNoN)) {
        SV_TRY(M.ke.ensure((size_t)M.nEl * 16 * dof * dof * sizeof(double)));
        SV_TRY(M.re.ensure((size_t)M.nEl * 4 * dof * sizeof(double)));
        const unsigned grid = (unsigned)(((size_t)M.nEl * 16 + 255) / 256);
        if (phys == SVFSI_PHYS_LELAS)
            hipLaunchKernelGGL(k_asm_lelas, dim3(grid), dim3(256), 0, s, 0, M.nEl, M.ien.as<int>(),
                               M.elemId.as<int>(), M.emap.as<int>(), M.x.as<double>(), tDof, Ag, Dg, prop, P, rule,
                               val, R, M.ke.as<double>(), M.re.as<double>(), M.gSlot.as<int>(), M.rSlot.as<int>());
        else
            hipLaunchKernelGGL(k_asm_fluid, dim3(grid), dim3(256), 0, s, 0, M.nEl, M.ien.as<int>(),
                               M.elemId.as<int>(), M.emap.as<int>(), M.x.as<double>(), tDof, Ag, Yg, prop, P, rule,
                               val, R, M.ke.as<double>(), M.re.as<double>(), M.gSlot.as<int>(), M.rSlot.as<int>());
        SV_HIP(hipGetLastError());
        return launch_gather(M, dof, val, R, s);
    }
    const int nC = (int)M.colorOff.size() - 1;
    for (int c = 0; c < nC; ++c) {
        const int begin = M.colorOff[c], count = M.colorOff[c + 1] - begin;
        if (count == 0) continue;
        const unsigned grid = (unsigned)((count * 16 + 255) / 256);
        if (phys == SVFSI_PHYS_LELAS)
            hipLaunchKernelGGL(k_asm_lelas, dim3(grid), dim3(256), 0, s, begin, count, M.ien.as<int>(),
                               M.elemId.as<int>(), M.emap.as<int>(), M.x.as<double>(), tDof, Ag, Dg, prop,
                               P, rule, val, R, nullptr, nullptr, nullptr, nullptr);
        else
            hipLaunchKernelGGL(k_asm_fluid, dim3(grid), dim3(256), 0, s, begin, count, M.ien.as<int>(),
                               M.elemId.as<int>(), M.emap.as<int>(), M.x.as<double>(), tDof, Ag, Yg, prop,
                               P, rule, val, R, nullptr, nullptr, nullptr, nullptr);
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

int launch_face_traction(const Mesh& M, const Mesh::FaceSet& F, int dof, const double* h,
                         const Tri3Rule& rule, double* R, hipStream_t s)
{
    const int nC = (int)F.colorOff.size() - 1;
    for (int c = 0; c < nC; ++c) {
        const int begin = F.colorOff[c], count = F.colorOff[c + 1] - begin;
        if (count == 0) continue;
        hipLaunchKernelGGL(k_face_traction, dim3((count * 3 + 255) / 256), dim3(256), 0, s, begin, count,
                           F.order.as<int>(), F.fien.as<int>(), F.inner.as<int>(), M.x.as<double>(), dof,
                           h[0], h[1], h[2], rule, R);
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

int launch_face_backflow(const Mesh& M, const Mesh::FaceSet& F, const AsmParams& P, int tDof,
                         const double* Yg, const Tri3Rule& rule, double* val, double* R, hipStream_t s)
{
    const int nC = (int)F.colorOff.size() - 1;
    for (int c = 0; c < nC; ++c) {
        const int begin = F.colorOff[c], count = F.colorOff[c + 1] - begin;
        if (count == 0) continue;
        hipLaunchKernelGGL(k_face_backflow, dim3((count * 9 + 255) / 256), dim3(256), 0, s, begin, count,
                           F.order.as<int>(), F.fien.as<int>(), F.inner.as<int>(), F.fmap.as<int>(),
                           M.x.as<double>(), tDof, Yg, P, rule, val, R);
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

int launch_fluid_dmu(const Mesh& M, const AsmParams& P, const Tet4Rule& rule, int tDof, const double* Ag,
                     const double* Yg, const double* muEl, const double* lam, const double* du, double* g,
                     hipStream_t s)
{
    if (M.nEl == 0) return SVFSI_OK;
    const unsigned grid = (unsigned)(((size_t)M.nEl * 16 + 255) / 256);
    hipLaunchKernelGGL(k_fluid_dmu, dim3(grid), dim3(256), 0, s, M.nEl, M.ien.as<int>(), M.elemId.as<int>(),
                       M.x.as<double>(), tDof, Ag, Yg, muEl, P, rule, lam, du, g);
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

int launch_lelas_dkde(const Mesh& M, const AsmParams& P, const Tet4Rule& rule, const double* lam,
                      const double* u, double* g, hipStream_t s)
{
    if (M.nEl == 0) return SVFSI_OK;
    hipLaunchKernelGGL(k_lelas_dkde, dim3((M.nEl + 255) / 256), dim3(256), 0, s, M.nEl, M.ien.as<int>(),
                       M.elemId.as<int>(), M.x.as<double>(), P, rule, lam, u, g);
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

int launch_asm_struct(Mesh& M, const AsmParams& P, int nG, const double* rule, int tDof, const double* Ag,
                      const double* Dg, double* val, double* R, int kern, hipStream_t s)
{
    if (nG != 4 && nG != 15) {
        set_error("STRUCT: nG %d not supported", nG);
        return SVFSI_ERR_NOTSUP;
    }
    if (kern == SVFSI_STRUCT_AUTO) kern = SVFSI_STRUCT_MFMA64;
    if (kern != SVFSI_STRUCT_VALU && kern != SVFSI_STRUCT_MFMA64 && kern != SVFSI_STRUCT_MFMA32) {
        set_error("STRUCT: unknown kernel selector %d", kern);
        return SVFSI_ERR_ARG;
    }
    const int nC = (int)M.colorOff.size() - 1;
    if (kern != SVFSI_STRUCT_VALU) {
        const size_t nPts = (size_t)M.nEl * nG;
        SV_TRY(M.srec.ensure(nPts * SR_HBM * sizeof(double)));
        double* rec = M.srec.as<double>();
        const unsigned gp = (unsigned)((nPts + 255) / 256);
        {
            ProfScope ps(3, (double)nPts * SR_HBM * 8.0);
            if (nG == 4)
                SV_LAUNCH(ps, (k_struct_points<4>), dim3(gp), dim3(256), 0, s, M.nEl, M.ien.as<int>(),
                          M.x.as<double>(), tDof, Ag, Dg, P, rule, rec);
            else
                SV_LAUNCH(ps, (k_struct_points<15>), dim3(gp), dim3(256), 0, s, M.nEl, M.ien.as<int>(),
                          M.x.as<double>(), tDof, Ag, Dg, P, rule, rec);
        }
        SV_HIP(hipGetLastError());
        constexpr int EPW4 = 4, EPW15 = 2;
        const bool gather = gather_enabled(M.eNoN);
        double* ke = nullptr;
        double* re = nullptr;
        if (gather) {
            SV_TRY(M.ke.ensure((size_t)M.nEl * 100 * 9 * sizeof(double)));
            SV_TRY(M.re.ensure((size_t)M.nEl * 10 * 3 * sizeof(double)));
            ke = M.ke.as<double>();
            re = M.re.as<double>();
        }
        // gather: one launch over all elements; scatter: one per colour
        const int nL = gather ? 1 : nC;
        for (int c = 0; c < nL; ++c) {
            const int begin = gather ? 0 : M.colorOff[c];
            const int count = gather ? M.nEl : M.colorOff[c + 1] - begin;
            if (count == 0) continue;
            const int epw = nG == 4 ? EPW4 : EPW15;
            const dim3 grid((count + epw - 1) / epw);
            ProfScope ps(2, 2.0 * 30 * 30 * 6 * nG * (double)count);
#define SV_SMFMA(NGV, MT, EPWV)                                                                                     \
    SV_LAUNCH(ps, (k_asm_struct_mfma<NGV, MT, EPWV>), grid, dim3(256), 0, s, begin, count, M.ien.as<int>(),   \
              M.emap.as<int>(), P, rule, rec, val, R, ke, re, M.gSlot.as<int>(), M.rSlot.as<int>())
            if (nG == 4) {
                if (kern == SVFSI_STRUCT_MFMA64) SV_SMFMA(4, double, EPW4);
                else SV_SMFMA(4, float, EPW4);
            } else {
                if (kern == SVFSI_STRUCT_MFMA64) SV_SMFMA(15, double, EPW15);
                else SV_SMFMA(15, float, EPW15);
            }
#undef SV_SMFMA
        }
        SV_HIP(hipGetLastError());
        return gather ? launch_gather(M, 3, val, R, s) : SVFSI_OK;
    }
    for (int c = 0; c < nC; ++c) {
        const int begin = M.colorOff[c], count = M.colorOff[c + 1] - begin;
        if (count == 0) continue;
        ProfScope ps(2, 2.0 * 30 * 30 * 6 * nG * (double)count);
        if (nG == 4) {
            constexpr int EPW = 8;
            SV_LAUNCH(ps, (k_asm_struct<4, EPW>), dim3((count + EPW - 1) / EPW), dim3(256), 0, s, begin, count,
                      M.ien.as<int>(), M.emap.as<int>(), M.x.as<double>(), tDof, Ag, Dg, P, rule, val, R);
        } else {
            constexpr int EPW = 2;
            SV_LAUNCH(ps, (k_asm_struct<15, EPW>), dim3((count + EPW - 1) / EPW), dim3(256), 0, s, begin, count,
                      M.ien.as<int>(), M.emap.as<int>(), M.x.as<double>(), tDof, Ag, Dg, P, rule, val, R);
        }
    }
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

} // namespace svfsi
