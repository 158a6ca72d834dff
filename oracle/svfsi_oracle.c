/*
 * svfsi_oracle.c -- CPU restatement of the svFSI per-Newton-step
 * linear-system path (element kernels, DOASSEM, svFSILS Krylov solve).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product library.
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg.  See svfsi_oracle.h for layouts.  Citations are
 * file:line under Code/Source/ of SimVascular/svFSI.
 */
#define _POSIX_C_SOURCE 199309L /* clock_gettime under -std=c11 */
#include "svfsi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <time.h>
#include <complex.h>

#define EPS DBL_EPSILON

/* ===================================================================== */
/* NN.f                                                                   */
/* ===================================================================== */

/* Threads for the solver kernels below (SpMV, dots, vector updates).  The
 * default 1 runs them serially, in the order of the reference's one-rank
 * loops, and the tests use that.  More threads are used only by the CPU
 * baseline timing, as a stand-in for the reference's MPI ranks on the
 * host cores.  With one thread an OpenMP reduction adds its partial sum to
 * 0.0, so the results are bitwise those of the plain loops. */
static int orc_nthreads = 1;
void orc_set_threads(int n) { orc_nthreads = n < 1 ? 1 : n; }
#define ORC_PAR _Pragma("omp parallel for if(orc_nthreads > 1) num_threads(orc_nthreads) schedule(static)")
#define ORC_PAR_SUM(v) _Pragma(ORC_STR(omp parallel for if(orc_nthreads > 1) num_threads(orc_nthreads) schedule(static) reduction(+:v)))
#define ORC_STR(x) #x

/* ISZERO (svFSI/MATFUN.f:763-779, UTIL.f:879-903 with ib absent):
 * true iff |a| / max(|a|, eps) < 10 eps. */
int orc_iszero(double a)
{
    double aa = fabs(a);
    double nrm = aa > EPS ? aa : EPS;
    return (aa / nrm) < 10.0 * EPS;
}

/* GETGIP, TET4 (svFSI/NN.f:266-275). */
void orc_rule_tet4(double w[4], double xi[4][3])
{
    double s = (5.0 + 3.0 * sqrt(5.0)) / 20.0;
    double t = (5.0 - sqrt(5.0)) / 20.0;
    for (int g = 0; g < 4; ++g) {
        w[g] = 1.0 / 24.0;
        for (int i = 0; i < 3; ++i) xi[g][i] = t;
        if (g < 3) xi[g][g] = s;
    }
}

/* Four-point rule for TET10 (the configuration's reduced rule; the
 * reference always uses 15 points for TET10, svFSI/NN.f:71-76).  Same
 * points and weights as the TET4 rule; exact for quadratic integrands. */
void orc_rule_tet10_4pt(double w[4], double xi[4][3]) { orc_rule_tet4(w, xi); }

/* GETGIP, TET10 15-point rule (svFSI/NN.f:277-307). */
void orc_rule_tet10(double w[15], double xi[15][3])
{
    w[0] = 0.030283678097089;
    for (int g = 1; g < 5; ++g) w[g] = 0.006026785714286;
    for (int g = 5; g < 9; ++g) w[g] = 0.011645249086029;
    for (int g = 9; g < 15; ++g) w[g] = 0.010949141561386;
    double s = 0.25;
    xi[0][0] = s; xi[0][1] = s; xi[0][2] = s;
    s = 0.333333333333333; double t = 0.0;
    double p1[4][3] = {{t, s, s}, {s, t, s}, {s, s, t}, {s, s, s}};
    for (int g = 0; g < 4; ++g) for (int i = 0; i < 3; ++i) xi[1 + g][i] = p1[g][i];
    s = 0.090909090909091; t = 0.727272727272727;
    double p2[4][3] = {{t, s, s}, {s, t, s}, {s, s, t}, {s, s, s}};
    for (int g = 0; g < 4; ++g) for (int i = 0; i < 3; ++i) xi[5 + g][i] = p2[g][i];
    s = 0.066550153573664; t = 0.433449846426336;
    double p3[6][3] = {{s, s, t}, {s, t, s}, {s, t, t}, {t, t, s}, {t, s, t}, {t, s, s}};
    for (int g = 0; g < 6; ++g) for (int i = 0; i < 3; ++i) xi[9 + g][i] = p3[g][i];
}

/* GETGIP, TRI3 (svFSI/NN.f:416-422). */
void orc_rule_tri3(double w[3], double xi[3][2])
{
    double s = 2.0 / 3.0, t = 1.0 / 6.0;
    w[0] = w[1] = w[2] = 1.0 / 6.0;
    xi[0][0] = t; xi[0][1] = t;
    xi[1][0] = s; xi[1][1] = t;
    xi[2][0] = t; xi[2][1] = s;
}

/* GETGNN, TET4 (svFSI/NN.f:655-671). */
void orc_shape_tet4(const double xi[3], double N[4], double Nxi[4][3])
{
    N[0] = xi[0]; N[1] = xi[1]; N[2] = xi[2];
    N[3] = 1.0 - xi[0] - xi[1] - xi[2];
    memset(Nxi, 0, sizeof(double) * 12);
    Nxi[0][0] = 1.0; Nxi[1][1] = 1.0; Nxi[2][2] = 1.0;
    Nxi[3][0] = Nxi[3][1] = Nxi[3][2] = -1.0;
}

/* GETGNN, TET10 (svFSI/NN.f:674-715).  Mid-edge order:
 * 5:(1,2) 6:(2,3) 7:(1,3) 8:(1,4) 9:(2,4) 10:(3,4). */
void orc_shape_tet10(const double xi[3], double N[10], double Nxi[10][3])
{
    double x = xi[0], y = xi[1], z = xi[2];
    double s = 1.0 - x - y - z;
    N[0] = x * (2.0 * x - 1.0);
    N[1] = y * (2.0 * y - 1.0);
    N[2] = z * (2.0 * z - 1.0);
    N[3] = s * (2.0 * s - 1.0);
    N[4] = 4.0 * x * y;
    N[5] = 4.0 * y * z;
    N[6] = 4.0 * x * z;
    N[7] = 4.0 * x * s;
    N[8] = 4.0 * y * s;
    N[9] = 4.0 * z * s;
    double d[10][3] = {
        {4.0 * x - 1.0, 0.0, 0.0},
        {0.0, 4.0 * y - 1.0, 0.0},
        {0.0, 0.0, 4.0 * z - 1.0},
        {1.0 - 4.0 * s, 1.0 - 4.0 * s, 1.0 - 4.0 * s},
        {4.0 * y, 4.0 * x, 0.0},
        {0.0, 4.0 * z, 4.0 * y},
        {4.0 * z, 0.0, 4.0 * x},
        {4.0 * (s - x), -4.0 * x, -4.0 * x},
        {-4.0 * y, 4.0 * (s - y), -4.0 * y},
        {-4.0 * z, -4.0 * z, 4.0 * (s - z)}};
    memcpy(Nxi, d, sizeof(d));
}

/* GETGNN, TRI3 (svFSI/NN.f:1100-1110). */
void orc_shape_tri3(const double xi[2], double N[3], double Nxi[3][2])
{
    N[0] = xi[0]; N[1] = xi[1]; N[2] = 1.0 - xi[0] - xi[1];
    Nxi[0][0] = 1.0; Nxi[0][1] = 0.0;
    Nxi[1][0] = 0.0; Nxi[1][1] = 1.0;
    Nxi[2][0] = -1.0; Nxi[2][1] = -1.0;
}

/* GNN, insd = 3 (svFSI/NN.f:1469-1565): isoparametric Jacobian, its
 * cofactor inverse, metric ks = xiX^T xiX and physical gradients. */
void orc_gnn(int eNoN, const double (*Nxi)[3], const double (*xl)[3],
             double (*Nx)[3], double* Jac, double ks[3][3])
{
    double A[3][3] = {{0}}; /* A[i][j] = xXi(i,j) = sum_a x_i(a) dN_a/dxi_j */
    for (int a = 0; a < eNoN; ++a)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) A[i][j] += xl[a][i] * Nxi[a][j];
    double J = A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0]
             + A[0][2] * A[1][0] * A[2][1] - A[0][0] * A[1][2] * A[2][1]
             - A[0][1] * A[1][0] * A[2][2] - A[0][2] * A[1][1] * A[2][0];
    double B[3][3]; /* B[i][j] = xiX(i,j) = d xi_i / d x_j */
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
        for (int j = 0; j < 3; ++j)
            ks[i][j] = B[0][i] * B[0][j] + B[1][i] * B[1][j] + B[2][i] * B[2][j];
    for (int a = 0; a < eNoN; ++a)
        for (int j = 0; j < 3; ++j)
            Nx[a][j] = Nxi[a][0] * B[0][j] + Nxi[a][1] * B[1][j] + Nxi[a][2] * B[2][j];
    *Jac = J;
}

/* ===================================================================== */
/* LHSA.f                                                                 */
/* ===================================================================== */

/* LHSA / ADDCOL (svFSI/LHSA.f:38-262): for every element, every (a,b)
 * node pair inserts column IEN(b) in row IEN(a), kept sorted and unique by
 * insertion sort.  Pass 1 counts per-row lengths into rowPtr and returns
 * nnz; pass 2 (orc_lhsa_fill) writes the columns. */
static int* g_uind = NULL;   /* [nNo][cap] insertion lists */
static int* g_ulen = NULL;
static int  g_cap = 0;

static void addcol(int row, int col)
{
    int* u = g_uind + (size_t)row * g_cap;
    int n = g_ulen[row];
    int i = 0;
    while (i < n && u[i] < col) ++i;
    if (i < n && u[i] == col) return;
    if (n == g_cap) abort(); /* capacity chosen large enough below */
    for (int j = n; j > i; --j) u[j] = u[j - 1];
    u[i] = col;
    g_ulen[row] = n + 1;
}

static void lhsa_build(int nNo, int nEl, int eNoN, const int* IEN)
{
    /* node-to-element degree bounds the row length: <= deg*eNoN */
    int* deg = calloc((size_t)nNo, sizeof(int));
    for (size_t k = 0; k < (size_t)nEl * eNoN; ++k) deg[IEN[k]]++;
    int mx = 0;
    for (int a = 0; a < nNo; ++a) if (deg[a] > mx) mx = deg[a];
    free(deg);
    g_cap = mx * eNoN + 1;
    free(g_uind); free(g_ulen);
    g_uind = malloc(sizeof(int) * (size_t)nNo * g_cap);
    g_ulen = calloc((size_t)nNo, sizeof(int));
    for (int e = 0; e < nEl; ++e)
        for (int a = 0; a < eNoN; ++a)
            for (int b = 0; b < eNoN; ++b)
                addcol(IEN[(size_t)e * eNoN + a], IEN[(size_t)e * eNoN + b]);
}

int orc_lhsa_count(int nNo, int nEl, int eNoN, const int* IEN, int* rowPtr)
{
    lhsa_build(nNo, nEl, eNoN, IEN);
    rowPtr[0] = 0;
    for (int r = 0; r < nNo; ++r) {
        if (g_ulen[r] == 0) return -1 - r; /* "Node is isolated" LHSA.f:177 */
        rowPtr[r + 1] = rowPtr[r] + g_ulen[r];
    }
    return rowPtr[nNo];
}

void orc_lhsa_fill(int nNo, int nEl, int eNoN, const int* IEN,
                   const int* rowPtr, int* colPtr)
{
    (void)nEl; (void)eNoN; (void)IEN;
    for (int r = 0; r < nNo; ++r)
        memcpy(colPtr + rowPtr[r], g_uind + (size_t)r * g_cap, sizeof(int) * g_ulen[r]);
    free(g_uind); free(g_ulen);
    g_uind = NULL; g_ulen = NULL;
}

/* DOASSEM (svFSI/LHSA.f:266-298): R(:,row) += lR(:,a); per (a,b) a
 * bisection for the column inside the row, then Val(:,ptr) += lK(:,a,b). */
void orc_doassem(int d, const int* eqN, int dof, const double* lK,
                 const double* lR, const int* rowPtr, const int* colPtr,
                 double* Val, double* R)
{
    int dd = dof * dof;
    for (int a = 0; a < d; ++a) {
        int rowN = eqN[a];
        for (int i = 0; i < dof; ++i) R[(size_t)rowN * dof + i] += lR[a * dof + i];
        for (int b = 0; b < d; ++b) {
            int colN = eqN[b];
            int left = rowPtr[rowN], right = rowPtr[rowN + 1];
            int ptr = (left + right) / 2;
            while (colN != colPtr[ptr]) {
                if (colN > colPtr[ptr]) left = ptr; else right = ptr;
                ptr = (left + right) / 2;
            }
            const double* k = lK + ((size_t)b * d + a) * dd;
            double* v = Val + (size_t)ptr * dd;
            for (int q = 0; q < dd; ++q) v[q] += k[q];
        }
    }
}

/* ===================================================================== */
/* LELAS.f                                                                */
/* ===================================================================== */

/* LELAS3D (svFSI/LELAS.f:130-242), no prestress, body-force field Bf = 0,
 * elasticity modulus passed per element. */
void orc_lelas3d(int eNoN, double w, const double* N, const double (*Nx)[3],
                 const double* al, const double* dl, int tDof,
                 const orc_eq* eq, double E, double* lR, double* lK)
{
    const int dof = 3, dd = 9;
    double rho = eq->rho, nu = eq->nu;
    double lambda = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu);
    double mu = E * 0.5 / (1.0 + nu);
    double lDm = lambda / mu;
    double T1 = eq->af * eq->beta * eq->dt * eq->dt;
    double amd = eq->am / T1 * rho;
    double wl = w * T1 * mu;
    double ud[3] = {-eq->fb[0], -eq->fb[1], -eq->fb[2]};
    double ed[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < eNoN; ++a) {
        const double* A = al + a * tDof;
        const double* D = dl + a * tDof;
        for (int i = 0; i < 3; ++i) ud[i] += N[a] * A[i];
        ed[0] += Nx[a][0] * D[0];
        ed[1] += Nx[a][1] * D[1];
        ed[2] += Nx[a][2] * D[2];
        ed[3] += Nx[a][1] * D[0] + Nx[a][0] * D[1];
        ed[4] += Nx[a][2] * D[1] + Nx[a][1] * D[2];
        ed[5] += Nx[a][0] * D[2] + Nx[a][2] * D[0];
    }
    double divD = lambda * (ed[0] + ed[1] + ed[2]);
    double S[6];
    S[0] = divD + 2.0 * mu * ed[0];
    S[1] = divD + 2.0 * mu * ed[1];
    S[2] = divD + 2.0 * mu * ed[2];
    S[3] = mu * ed[3];
    S[4] = mu * ed[4];
    S[5] = mu * ed[5];
    for (int a = 0; a < eNoN; ++a) {
        lR[a * dof + 0] += w * (rho * N[a] * ud[0] + Nx[a][0] * S[0] + Nx[a][1] * S[3] + Nx[a][2] * S[5]);
        lR[a * dof + 1] += w * (rho * N[a] * ud[1] + Nx[a][0] * S[3] + Nx[a][1] * S[1] + Nx[a][2] * S[4]);
        lR[a * dof + 2] += w * (rho * N[a] * ud[2] + Nx[a][0] * S[5] + Nx[a][1] * S[4] + Nx[a][2] * S[2]);
        for (int b = 0; b < eNoN; ++b) {
            double* k = lK + ((size_t)b * eNoN + a) * dd;
            double NxdNx = Nx[a][0] * Nx[b][0] + Nx[a][1] * Nx[b][1] + Nx[a][2] * Nx[b][2];
            double T = amd * N[a] * N[b] / mu + NxdNx;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double v = lDm * Nx[a][i] * Nx[b][j] + Nx[a][j] * Nx[b][i];
                    if (i == j) v = T + (1.0 + lDm) * Nx[a][i] * Nx[b][i];
                    k[i * dof + j] += wl * v;
                }
        }
    }
}

/* CONSTRUCT_LELAS (svFSI/LELAS.f:38-128): tet4 only, GNN once (lShpF). */
void orc_construct_lelas(int nNo, int nEl, const int* IEN, const double* x,
                         const double* Ag, const double* Dg, int tDof,
                         const orc_eq* eq, const double* E_el,
                         const int* rowPtr, const int* colPtr,
                         double* Val, double* R)
{
    (void)nNo;
    const int eNoN = 4, dof = 3;
    double wg[4], xi[4][3], Ng[4][4], Nxi[4][3];
    orc_rule_tet4(wg, xi);
    for (int g = 0; g < 4; ++g) orc_shape_tet4(xi[g], Ng[g], Nxi);
    double xl[4][3], al[4 * 3], dl[4 * 3], Nx[4][3], ks[3][3], Jac = 0.0;
    double lR[4 * 3], lK[16 * 9];
    int ptr[4];
    for (int e = 0; e < nEl; ++e) {
        for (int a = 0; a < eNoN; ++a) {
            int Ac = IEN[(size_t)e * eNoN + a];
            ptr[a] = Ac;
            for (int i = 0; i < 3; ++i) {
                xl[a][i] = x[(size_t)Ac * 3 + i];
                al[a * 3 + i] = Ag[(size_t)Ac * tDof + i];
                dl[a * 3 + i] = Dg[(size_t)Ac * tDof + i];
            }
        }
        memset(lR, 0, sizeof(lR));
        memset(lK, 0, sizeof(lK));
        double E = E_el ? E_el[e] : eq->E;
        for (int g = 0; g < 4; ++g) {
            if (g == 0) orc_gnn(eNoN, (const double (*)[3])Nxi, (const double (*)[3])xl, Nx, &Jac, ks);
            double w = wg[g] * Jac;
            orc_lelas3d(eNoN, w, Ng[g], (const double (*)[3])Nx, al, dl, 3, eq, E, lR, lK);
        }
        orc_doassem(eNoN, ptr, dof, lK, lR, rowPtr, colPtr, Val, R);
    }
}

/* ===================================================================== */
/* FLUID.f                                                                */
/* ===================================================================== */

/* GETVISCOSITY (svFSI/FLUID.f:1690-1739): viscosity mu and d mu / d gamma
 * (mu_x) at shear rate *gamma for the model eq->viscType, mu_i = mu_in.
 * Casson raises *gamma to the threshold lam below it (gamma is INOUT in
 * the reference, and the caller divides by the raised value). */
static void orc_getviscosity_d(const orc_eq* eq, double mu_in, double* gamma, double* mu, double* mu_s,
                               double* mu_x)
{
    const int vt = eq->viscType == 0 ? 698 : eq->viscType;
    if (vt == 697) {
        const double mu_i = mu_in, mu_o = eq->mu_o, lam = eq->lam, a = eq->visc_a, n = eq->visc_n;
        double T1 = 1.0 + pow(lam * *gamma, a);
        double T2 = pow(T1, (n - 1.0) / a);
        *mu = mu_i + (mu_o - mu_i) * T2;
        *mu_s = mu_i;
        T1 = T2 / T1;
        T2 = pow(lam, a) * pow(*gamma, a - 1.0) * T1;
        *mu_x = (mu_o - mu_i) * (n - 1.0) * T2;
    } else if (vt == 696) {
        const double mu_i = mu_in, lam = eq->lam;
        double mu_o = eq->mu_o;
        if (*gamma < lam) {
            mu_o = mu_o / sqrt(lam);
            *gamma = lam;
        } else {
            mu_o = mu_o / sqrt(*gamma);
        }
        *mu = (mu_i + mu_o) * (mu_i + mu_o);
        *mu_s = mu_i * mu_i;
        *mu_x = 2.0 * mu_o * (mu_o + mu_i) / *gamma;
    } else {
        *mu = mu_in;
        *mu_s = mu_in;
        *mu_x = 0.0;
    }
}

void orc_getviscosity(const orc_eq* eq, double mu_i, double gamma_in, double* gamma_out, double* mu,
                      double* mu_s, double* mu_x)
{
    double g = gamma_in;
    orc_getviscosity_d(eq, mu_i, &g, mu, mu_s, mu_x);
    if (gamma_out) *gamma_out = g;
}

#define FREAL double
#define FSQRT sqrt
#define FRE(x) (x)
#define FN(n) n
#define FSTATIC
#include "fluid_elem.inc"
#undef FREAL
#undef FSQRT
#undef FRE
#undef FN
#undef FSTATIC

/* CONSTRUCT_FLUID (svFSI/FLUID.f:40-190), tet4 VMS (nFs = 1): both
 * function spaces are the mesh tet4 basis (GETTHOODFS, FS.f:257-286);
 * GNN at g = 1 only (lShpF); GNNxx of an affine tet is zero. */
void orc_construct_fluid(int nNo, int nEl, const int* IEN, const double* x,
                         const double* Ag, const double* Yg, int tDof,
                         const orc_eq* eq, const double* mu_el,
                         const int* rowPtr, const int* colPtr,
                         double* Val, double* R)
{
    (void)nNo;
    const int eNoN = 4, dof = 4;
    double wg[4], xi[4][3], Ng[4][4], Nxi[4][3];
    orc_rule_tet4(wg, xi);
    for (int g = 0; g < 4; ++g) orc_shape_tet4(xi[g], Ng[g], Nxi);
    double Nxx[4][6];
    memset(Nxx, 0, sizeof(Nxx));
    double xl[4][3], al[16], yl[16], Nx[4][3], ks[3][3], Jac = 0.0;
    double lR[16], lK[16 * 16];
    int ptr[4];
    for (int e = 0; e < nEl; ++e) {
        for (int a = 0; a < eNoN; ++a) {
            int Ac = IEN[(size_t)e * eNoN + a];
            ptr[a] = Ac;
            for (int i = 0; i < 3; ++i) xl[a][i] = x[(size_t)Ac * 3 + i];
            for (int i = 0; i < 4; ++i) {
                al[a * 4 + i] = Ag[(size_t)Ac * tDof + i];
                yl[a * 4 + i] = Yg[(size_t)Ac * tDof + i];
            }
        }
        memset(lR, 0, sizeof(lR));
        memset(lK, 0, sizeof(lK));
        double mu = mu_el ? mu_el[e] : eq->mu;
        orc_gnn(eNoN, (const double (*)[3])Nxi, (const double (*)[3])xl, Nx, &Jac, ks);
        for (int g = 0; g < 4; ++g)
            orc_fluid3d_m(1, eNoN, eNoN, wg[g] * Jac, (const double (*)[3])ks, Ng[g], Ng[g],
                          (const double (*)[3])Nx, (const double (*)[3])Nx, (const double (*)[6])Nxx,
                          al, yl, 4, eq, mu, lR, lK);
        for (int g = 0; g < 4; ++g)
            orc_fluid3d_c(1, eNoN, eNoN, wg[g] * Jac, (const double (*)[3])ks, Ng[g], Ng[g],
                          (const double (*)[3])Nx, (const double (*)[3])Nx, (const double (*)[6])Nxx,
                          al, yl, 4, eq, mu, lR, lK);
        orc_doassem(eNoN, ptr, dof, lK, lR, rowPtr, colPtr, Val, R);
    }
}

/* GETGNNxx, TET10 (svFSI/NN.f:1285-1295): the constant second derivatives
 * d2N/dxi_i dxi_j in the order (11, 22, 33, 12, 23, 13). */
void orc_shape_tet10_xx(double Nxi2[10][6])
{
    static const double t[10][6] = {
        {4, 0, 0, 0, 0, 0}, {0, 4, 0, 0, 0, 0}, {0, 0, 4, 0, 0, 0}, {4, 4, 4, 4, 4, 4},
        {0, 0, 0, 4, 0, 0}, {0, 0, 0, 0, 4, 0}, {0, 0, 0, 0, 0, 4},
        {-8, 0, 0, -4, 0, -4}, {0, -8, 0, -4, -4, 0}, {0, 0, -8, 0, -4, -4}};
    memcpy(Nxi2, t, sizeof(t));
}

/* Solves K X = B for X (6 x 6 K, nrhs right-hand sides stored B[r][i]) by
 * LU factorisation with partial (row) pivoting, the algorithm of LAPACK's
 * DGESV that GNNxx calls (svFSI/NN.f:1672).  Returns 0, or 1 if K is
 * singular. */
static int solve6(double K[6][6], double (*B)[6], int nrhs)
{
    int piv[6];
    for (int k = 0; k < 6; ++k) {
        int p = k;
        for (int i = k + 1; i < 6; ++i)
            if (fabs(K[i][k]) > fabs(K[p][k])) p = i;
        piv[k] = p;
        if (K[p][k] == 0.0) return 1;
        if (p != k)
            for (int j = 0; j < 6; ++j) { double t = K[k][j]; K[k][j] = K[p][j]; K[p][j] = t; }
        for (int i = k + 1; i < 6; ++i) {
            K[i][k] /= K[k][k];
            for (int j = k + 1; j < 6; ++j) K[i][j] -= K[i][k] * K[k][j];
        }
    }
    for (int r = 0; r < nrhs; ++r) {
        double* b = B[r];
        for (int k = 0; k < 6; ++k)
            if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
        for (int i = 1; i < 6; ++i)
            for (int j = 0; j < i; ++j) b[i] -= K[i][j] * b[j];
        for (int i = 5; i >= 0; --i) {
            for (int j = i + 1; j < 6; ++j) b[i] -= K[i][j] * b[j];
            b[i] /= K[i][i];
        }
    }
    return 0;
}

/* GNNxx (svFSI/NN.f:1568-1678), 3-D: the physical second derivatives
 * Nxx(a) (order xx, yy, zz, xy, yz, xz) from the chain rule
 * d2N/dxi_i dxi_j = sum_kl x_k,i x_l,j N,kl + sum_k x_k,ij N,k, i.e. the
 * 6 x 6 system K Nxx(a) = Nxi2(a) - Nx(a) . xXi2 with K built from the
 * Jacobian xXi (NN.f:1632-1662).  Returns nonzero if K is singular. */
int orc_gnnxx(int eNoN, const double (*Nxi)[3], const double (*Nxi2)[6], const double (*xl)[3],
              const double (*Nx)[3], double (*Nxx)[6])
{
    double xXi[3][3] = {{0}}, xXi2[3][6] = {{0}};
    for (int a = 0; a < eNoN; ++a)
        for (int k = 0; k < 3; ++k) {
            for (int i = 0; i < 3; ++i) xXi[k][i] += xl[a][k] * Nxi[a][i];
            for (int i = 0; i < 6; ++i) xXi2[k][i] += xl[a][k] * Nxi2[a][i];
        }
    double K[6][6];
    for (int i = 0; i < 3; ++i) {
        K[i][0] = xXi[0][i] * xXi[0][i];
        K[i][1] = xXi[1][i] * xXi[1][i];
        K[i][2] = xXi[2][i] * xXi[2][i];
        K[i][3] = 2.0 * xXi[0][i] * xXi[1][i];
        K[i][4] = 2.0 * xXi[1][i] * xXi[2][i];
        K[i][5] = 2.0 * xXi[0][i] * xXi[2][i];
    }
    const int pr[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    for (int r = 0; r < 3; ++r) {
        const int i = pr[r][0], j = pr[r][1];
        double* k = K[3 + r];
        k[0] = xXi[0][i] * xXi[0][j];
        k[1] = xXi[1][i] * xXi[1][j];
        k[2] = xXi[2][i] * xXi[2][j];
        k[3] = xXi[0][i] * xXi[1][j] + xXi[0][j] * xXi[1][i];
        k[4] = xXi[1][i] * xXi[2][j] + xXi[1][j] * xXi[2][i];
        k[5] = xXi[0][i] * xXi[2][j] + xXi[0][j] * xXi[2][i];
    }
    for (int a = 0; a < eNoN; ++a)
        for (int i = 0; i < 6; ++i)
            Nxx[a][i] = Nxi2[a][i] - Nx[a][0] * xXi2[0][i] - Nx[a][1] * xXi2[1][i] - Nx[a][2] * xXi2[2][i];
    return solve6(K, Nxx, eNoN);
}

/* CONSTRUCT_FLUID (svFSI/FLUID.f:40-190) with a Taylor-Hood basis (nFs =
 * 2, P2P1): velocity on the tet10 mesh basis, pressure on the tet4 basis of
 * the corner nodes (SETTHOODFS, FS.f:201-255), vmsFlag = .FALSE.
 *   Integration 1 (GETTHOODFS iOpt 1, FS.f:288-308): the tet10 15-point
 *   rule; at every point GNN of the tet10 map gives Nwx, Jac and ksix and
 *   GNNxx gives Nwxx; the tet4 map's GNN (Nqx) at the first point only.
 *   Then FLUID3D_M.
 *   Integration 2 (iOpt 2, FS.f:310-330): the tet4 4-point rule; at every
 *   point GNN of the tet10 map (Nwx, Jac, ksix), then at the first point
 *   GNN of the tet4 map, whose Jac and ksix replace the tet10 ones there;
 *   Nwxx keeps its value from the last point of integration 1.  Then
 *   FLUID3D_C.
 * IEN[nEl][10], state (nNo, tDof).  Returns nonzero on a singular GNNxx. */
int orc_construct_fluid_th(int nNo, int nEl, const int* IEN, const double* x,
                           const double* Ag, const double* Yg, int tDof,
                           const orc_eq* eq, const double* mu_el,
                           const int* rowPtr, const int* colPtr,
                           double* Val, double* R)
{
    (void)nNo;
    enum { NW = 10, NQ = 4, dof = 4 };
    double w1[15], xi1[15][3], w2[4], xi2[4][3];
    orc_rule_tet10(w1, xi1);
    orc_rule_tet4(w2, xi2);
    double Nw1[15][NW], Nwxi1[15][NW][3], Nq1[15][NQ], Nqxi[NQ][3];
    double Nw2[4][NW], Nwxi2[4][NW][3], Nq2[4][NQ], Nxi2[NW][6];
    for (int g = 0; g < 15; ++g) {
        orc_shape_tet10(xi1[g], Nw1[g], Nwxi1[g]);
        orc_shape_tet4(xi1[g], Nq1[g], Nqxi);
    }
    for (int g = 0; g < 4; ++g) {
        orc_shape_tet10(xi2[g], Nw2[g], Nwxi2[g]);
        orc_shape_tet4(xi2[g], Nq2[g], Nqxi);
    }
    orc_shape_tet10_xx(Nxi2);
    double xl[NW][3], al[NW * 4], yl[NW * 4], Nwx[NW][3], Nqx[NQ][3], Nwxx[NW][6], ks[3][3], Jac = 0.0;
    double lR[NW * 4], lK[NW * NW * 16];
    int ptr[NW];
    for (int e = 0; e < nEl; ++e) {
        for (int a = 0; a < NW; ++a) {
            int Ac = IEN[(size_t)e * NW + a];
            ptr[a] = Ac;
            for (int i = 0; i < 3; ++i) xl[a][i] = x[(size_t)Ac * 3 + i];
            for (int i = 0; i < 4; ++i) {
                al[a * 4 + i] = Ag[(size_t)Ac * tDof + i];
                yl[a * 4 + i] = Yg[(size_t)Ac * tDof + i];
            }
        }
        memset(lR, 0, sizeof(lR));
        memset(lK, 0, sizeof(lK));
        const double mu = mu_el ? mu_el[e] : eq->mu;
        for (int g = 0; g < 15; ++g) {
            if (g == 0) orc_gnn(NQ, (const double (*)[3])Nqxi, (const double (*)[3])xl, Nqx, &Jac, ks);
            orc_gnn(NW, (const double (*)[3])Nwxi1[g], (const double (*)[3])xl, Nwx, &Jac, ks);
            if (orc_gnnxx(NW, (const double (*)[3])Nwxi1[g], (const double (*)[6])Nxi2, (const double (*)[3])xl,
                          (const double (*)[3])Nwx, Nwxx))
                return 1;
            orc_fluid3d_m(0, NW, NQ, w1[g] * Jac, (const double (*)[3])ks, Nw1[g], Nq1[g],
                          (const double (*)[3])Nwx, (const double (*)[3])Nqx, (const double (*)[6])Nwxx,
                          al, yl, 4, eq, mu, lR, lK);
        }
        for (int g = 0; g < 4; ++g) {
            orc_gnn(NW, (const double (*)[3])Nwxi2[g], (const double (*)[3])xl, Nwx, &Jac, ks);
            if (g == 0) orc_gnn(NQ, (const double (*)[3])Nqxi, (const double (*)[3])xl, Nqx, &Jac, ks);
            orc_fluid3d_c(0, NW, NQ, w2[g] * Jac, (const double (*)[3])ks, Nw2[g], Nq2[g],
                          (const double (*)[3])Nwx, (const double (*)[3])Nqx, (const double (*)[6])Nwxx,
                          al, yl, 4, eq, mu, lR, lK);
        }
        orc_doassem(NW, ptr, dof, lK, lR, rowPtr, colPtr, Val, R);
    }
    return 0;
}

/* THOOD_ValRC (svFSI/FS.f:352-407): every node that is a mid-edge node
 * (local 5..10) of a tet10 element has its continuity residual set to 0
 * and, in its block row, the pressure-pressure entry (4,4) set to 1 on the
 * diagonal block and to 0 elsewhere. */
void orc_thood_valrc(int nNo, int nEl, const int* IEN, const int* rowPtr, const int* colPtr, double* Val,
                     double* R)
{
    char* mid = (char*)calloc((size_t)nNo, 1);
    for (int e = 0; e < nEl; ++e)
        for (int a = 4; a < 10; ++a) mid[IEN[(size_t)e * 10 + a]] = 1;
    for (int a = 0; a < nNo; ++a) {
        if (!mid[a]) continue;
        R[(size_t)a * 4 + 3] = 0.0;
        for (int k = rowPtr[a]; k < rowPtr[a + 1]; ++k) Val[(size_t)k * 16 + 15] = colPtr[k] == a ? 1.0 : 0.0;
    }
    free(mid);
}

/* Complex-step instantiation of FLUID3D_M/C: with mu = mu_e + i h the
 * imaginary parts of lK, lR are h dK/dmu, h dR/dmu to rounding (no
 * subtractive cancellation), exact through tau_M, tau_C and tau_B. */
#define FREAL double complex
#define FSQRT csqrt
#define FRE(x) creal(x)
#define FN(n) n##_cs
#define FSTATIC static
#include "fluid_elem.inc"
#undef FREAL
#undef FSQRT
#undef FRE
#undef FN
#undef FSTATIC

/* Viscosity sensitivity of one linearised step (no reference counterpart):
 * g[e] = lam_e . (dR_e/dmu_e - dK_e/dmu_e du_e), the element term of
 * dL/dmu_e when K du = R and K^T lam = dL/du; element loop of
 * orc_construct_fluid with the derivative by complex step. */
void orc_fluid_dmu_dot(int nEl, const int* IEN, const double* x, const double* Ag,
                       const double* Yg, int tDof, const orc_eq* eq, const double* mu_el,
                       const double* lam, const double* du, double* g)
{
    const int eNoN = 4, dof = 4;
    double wg[4], xi[4][3], Ng[4][4], Nxi[4][3];
    orc_rule_tet4(wg, xi);
    for (int q = 0; q < 4; ++q) orc_shape_tet4(xi[q], Ng[q], Nxi);
    double Nxx[4][6];
    memset(Nxx, 0, sizeof(Nxx));
    double xl[4][3], al[16], yl[16], Nx[4][3], ks[3][3], Jac = 0.0;
    double complex lR[16], lK[16 * 16];
    int ptr[4];
    for (int e = 0; e < nEl; ++e) {
        for (int a = 0; a < eNoN; ++a) {
            int Ac = IEN[(size_t)e * eNoN + a];
            ptr[a] = Ac;
            for (int i = 0; i < 3; ++i) xl[a][i] = x[(size_t)Ac * 3 + i];
            for (int i = 0; i < 4; ++i) {
                al[a * 4 + i] = Ag[(size_t)Ac * tDof + i];
                yl[a * 4 + i] = Yg[(size_t)Ac * tDof + i];
            }
        }
        for (int q = 0; q < 16; ++q) lR[q] = 0.0;
        for (int q = 0; q < 256; ++q) lK[q] = 0.0;
        const double mu0 = mu_el ? mu_el[e] : eq->mu;
        const double h = 1e-30 * (mu0 != 0.0 ? fabs(mu0) : 1.0);
        const double complex mu = mu0 + h * I;
        orc_gnn(eNoN, (const double (*)[3])Nxi, (const double (*)[3])xl, Nx, &Jac, ks);
        for (int q = 0; q < 4; ++q)
            orc_fluid3d_m_cs(1, eNoN, eNoN, wg[q] * Jac, (const double (*)[3])ks, Ng[q], Ng[q],
                             (const double (*)[3])Nx, (const double (*)[3])Nx, (const double (*)[6])Nxx, al, yl,
                             4, eq, mu, lR, lK);
        for (int q = 0; q < 4; ++q)
            orc_fluid3d_c_cs(1, eNoN, eNoN, wg[q] * Jac, (const double (*)[3])ks, Ng[q], Ng[q],
                             (const double (*)[3])Nx, (const double (*)[3])Nx, (const double (*)[6])Nxx, al, yl,
                             4, eq, mu, lR, lK);
        double s = 0.0;
        for (int a = 0; a < eNoN; ++a)
            for (int r = 0; r < dof; ++r) {
                double t = cimag(lR[a * dof + r]) / h;
                for (int b = 0; b < eNoN; ++b)
                    for (int c = 0; c < dof; ++c)
                        t -= cimag(lK[((size_t)b * eNoN + a) * 16 + r * dof + c]) / h * du[(size_t)ptr[b] * dof + c];
                s += lam[(size_t)ptr[a] * dof + r] * t;
            }
        g[e] = s;
    }
}

/* ===================================================================== */
/* Face terms: GNNB + SETBCTRACL / BFLUID                                 */
/* ===================================================================== */

/* GNNB (svFSI/NN.f:1856-1996), tri3 face of a volume element: n is the
 * cross product of the two face tangents, flipped to point away from the
 * first element node that is not on the face.  Returns |n| via nV scale. */
static void gnnb_tri3(const int* fn, int Ec, const int* IEN, int eNoN,
                      const double* x, double n[3])
{
    const double* x1 = x + (size_t)fn[0] * 3;
    const double* x2 = x + (size_t)fn[1] * 3;
    const double* x3 = x + (size_t)fn[2] * 3;
    double t1[3], t2[3];
    for (int i = 0; i < 3; ++i) { t1[i] = x1[i] - x3[i]; t2[i] = x2[i] - x3[i]; }
    n[0] = t1[1] * t2[2] - t1[2] * t2[1];
    n[1] = t1[2] * t2[0] - t1[0] * t2[2];
    n[2] = t1[0] * t2[1] - t1[1] * t2[0];
    int inner = -1;
    for (int b = 0; b < eNoN && inner < 0; ++b) {
        int Bc = IEN[(size_t)Ec * eNoN + b];
        if (Bc != fn[0] && Bc != fn[1] && Bc != fn[2]) inner = Bc;
    }
    const double* xb = x + (size_t)inner * 3;
    double v0 = x1[0] - xb[0], v1 = x1[1] - xb[1], v2 = x1[2] - xb[2];
    if (n[0] * v0 + n[1] * v1 + n[2] * v2 < 0.0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
}

/* SETBCTRACL (svFSI/SETBC.f:319-416), steady uniform traction h. */
void orc_traction_tri3(int nFe, const int* fIEN, const int* fEl,
                       const int* IEN, int eNoN, const double* x,
                       const double h[3], int dof, const int* rowPtr,
                       const int* colPtr, double* Val, double* R)
{
    double wg[3], xi[3][2], N[3], Nxi[3][2];
    orc_rule_tri3(wg, xi);
    double lR[3 * 4], lK[9 * 16];
    for (int e = 0; e < nFe; ++e) {
        memset(lR, 0, sizeof(lR));
        memset(lK, 0, sizeof(lK));
        for (int g = 0; g < 3; ++g) {
            double n[3];
            gnnb_tri3(fIEN + 3 * e, fEl[e], IEN, eNoN, x, n);
            double Jac = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            double w = wg[g] * Jac;
            orc_shape_tri3(xi[g], N, Nxi);
            for (int a = 0; a < 3; ++a)
                for (int i = 0; i < 3; ++i) lR[a * dof + i] -= w * N[a] * h[i];
        }
        orc_doassem(3, fIEN + 3 * e, dof, lK, lR, rowPtr, colPtr, Val, R);
    }
}

/* BASSEMNEUBC (svFSI/EQASSEM.f:90-192) with h = 0 calling BFLUID
 * (svFSI/FLUID.f:1280-1336): outlet backflow stabilization. */
void orc_backflow_tri3(int nFe, const int* fIEN, const int* fEl,
                       const int* IEN, int eNoN, const double* x,
                       const double* Yg, int tDof, const orc_eq* eq,
                       const int* rowPtr, const int* colPtr,
                       double* Val, double* R)
{
    const int dof = 4, dd = 16;
    double wg[3], xi[3][2], N[3], Nxi[3][2];
    orc_rule_tri3(wg, xi);
    double lR[12], lK[9 * 16];
    for (int e = 0; e < nFe; ++e) {
        memset(lR, 0, sizeof(lR));
        memset(lK, 0, sizeof(lK));
        const int* fn = fIEN + 3 * e;
        for (int g = 0; g < 3; ++g) {
            double n[3];
            gnnb_tri3(fn, fEl[e], IEN, eNoN, x, n);
            double Jac = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            double nV[3] = {n[0] / Jac, n[1] / Jac, n[2] / Jac};
            double w = wg[g] * Jac;
            orc_shape_tri3(xi[g], N, Nxi);
            double y[3] = {0, 0, 0};
            for (int a = 0; a < 3; ++a)
                for (int i = 0; i < 3; ++i) y[i] += N[a] * Yg[(size_t)fn[a] * tDof + i];
            double wl = w * eq->af * eq->gam * eq->dt;
            double udn = y[0] * nV[0] + y[1] * nV[1] + y[2] * nV[2];
            udn = 0.5 * eq->bstab * eq->rho * (udn - fabs(udn));
            double hc[3] = {udn * y[0], udn * y[1], udn * y[2]};
            for (int a = 0; a < 3; ++a) {
                for (int i = 0; i < 3; ++i) lR[a * dof + i] -= w * N[a] * hc[i];
                for (int b = 0; b < 3; ++b) {
                    double T1 = wl * N[a] * N[b] * udn;
                    double* kk = lK + ((size_t)b * 3 + a) * dd;
                    kk[0] -= T1; kk[5] -= T1; kk[10] -= T1;
                }
            }
        }
        orc_doassem(3, fn, dof, lK, lR, rowPtr, colPtr, Val, R);
    }
}

/* ===================================================================== */
/* STRUCT.f + MATMODELS.f (neo-Hookean)                                   */
/* ===================================================================== */

static double det3(const double A[3][3])
{
    /* MAT_DET (svFSI/MATFUN.f:177-205): cofactor expansion on row 1 */
    return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1])
         - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
         + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}

static void inv3(const double A[3][3], double Ai[3][3])
{
    /* MAT_INV nd = 3 (svFSI/MATFUN.f:257-271) */
    double d = det3(A);
    Ai[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / d;
    Ai[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / d;
    Ai[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / d;
    Ai[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / d;
    Ai[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / d;
    Ai[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / d;
    Ai[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / d;
    Ai[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / d;
    Ai[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / d;
}

/* GETPK2CC, stIso_nHook branch without fibers or active stress
 * (svFSI/MATMODELS.f:41-150, 520-537), GETSVOLP (539-564), CCTOVOIGT
 * (1588-1647).  CC is formed as the reference does from TEN_DYADPROD and
 * TEN_SYMMPROD (svFSI/MATFUN.f:438-478). */
void orc_getpk2cc_nhk(const orc_eq* eq, const double F[3][3],
                      double S[3][3], double Dm[6][6])
{
    const double nd = 3.0;
    double Kp = eq->Kpen;
    double J = det3(F);
    double J2d = pow(J, -2.0 / nd);
    double C[3][3], Ci[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
    inv3(C, Ci);
    double p = 0.0, pl = 0.0;
    if (!orc_iszero(Kp)) {
        switch (eq->volType) {
        case 1: p = Kp * (J - 1.0); pl = Kp * (2.0 * J - 1.0); break;
        case 3: p = Kp * (1.0 - 1.0 / J); pl = Kp; break;
        default: p = 0.5 * Kp * (J - 1.0 / J); pl = Kp * J; break;
        }
    }
    double g1 = 2.0 * eq->C10;
    double Sb[3][3] = {{g1, 0, 0}, {0, g1, 0}, {0, 0, g1}};
    double CSb = 0.0;
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) CSb += C[i][j] * Sb[i][j];
    double r1 = J2d * CSb / nd;
    double Sd[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Sd[i][j] = J2d * Sb[i][j] - r1 * Ci[i][j];
    static double CC[3][3][3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                    double dy1 = Ci[i][j] * Sd[k][l] + Sd[i][j] * Ci[k][l];
                    double sym = 0.5 * (Ci[i][k] * Ci[j][l] + Ci[i][l] * Ci[j][k]);
                    double dyc = Ci[i][j] * Ci[k][l];
                    CC[i][j][k][l] = (-2.0 / nd) * dy1 + 2.0 * (r1 - p * J) * sym
                                   + (pl * J - 2.0 * r1 / nd) * dyc;
                }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S[i][j] = Sd[i][j] + p * J * Ci[i][j];
    /* Voigt order 11,22,33,12,23,31 */
    static const int vi[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {1, 2}, {2, 0}};
    for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) {
            Dm[a][b] = CC[vi[a][0]][vi[a][1]][vi[b][0]][vi[b][1]];
            Dm[b][a] = Dm[a][b];
        }
}

/* STRUCT3D (svFSI/STRUCT.f:164-496) for solid viscosity = 0, damping = 0,
 * no prestress and no fibers (the viscous Tv terms vanish). */
void orc_struct3d(int eNoN, double w, const double* N, const double (*Nx)[3],
                  const double* al, const double* yl, const double* dl,
                  int tDof, const orc_eq* eq, double* lR, double* lK)
{
    (void)yl;
    const int dof = 3, dd = 9;
    double rho = eq->rho;
    double afu = eq->af * eq->beta * eq->dt * eq->dt;
    double amd = eq->am * rho;
    double ud[3], F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i) ud[i] = -rho * eq->fb[i];
    for (int a = 0; a < eNoN; ++a) {
        const double* A = al + a * tDof;
        const double* D = dl + a * tDof;
        for (int i = 0; i < 3; ++i) {
            ud[i] += N[a] * (rho * A[i]);
            for (int j = 0; j < 3; ++j) F[i][j] += Nx[a][j] * D[i];
        }
    }
    double S[3][3], Dm[6][6], P[3][3];
    orc_getpk2cc_nhk(eq, (const double (*)[3])F, S, Dm);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            P[i][j] = F[i][0] * S[0][j] + F[i][1] * S[1][j] + F[i][2] * S[2][j];
    for (int a = 0; a < eNoN; ++a)
        for (int i = 0; i < 3; ++i)
            lR[a * dof + i] += w * (N[a] * ud[i] + Nx[a][0] * P[i][0] + Nx[a][1] * P[i][1] + Nx[a][2] * P[i][2]);
    /* Bm(6,3,a) (STRUCT.f:295-319) */
    double Bm[64][6][3];
    for (int a = 0; a < eNoN; ++a)
        for (int c = 0; c < 3; ++c) {
            Bm[a][0][c] = Nx[a][0] * F[c][0];
            Bm[a][1][c] = Nx[a][1] * F[c][1];
            Bm[a][2][c] = Nx[a][2] * F[c][2];
            Bm[a][3][c] = Nx[a][0] * F[c][1] + F[c][0] * Nx[a][1];
            Bm[a][4][c] = Nx[a][1] * F[c][2] + F[c][1] * Nx[a][2];
            Bm[a][5][c] = Nx[a][2] * F[c][0] + F[c][2] * Nx[a][0];
        }
    for (int b = 0; b < eNoN; ++b)
        for (int a = 0; a < eNoN; ++a) {
            double NxSNx = 0.0;
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) NxSNx += Nx[a][i] * S[i][j] * Nx[b][j];
            double T1 = amd * N[a] * N[b] + afu * NxSNx;
            double DBm[6][3];
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 3; ++c) {
                    double s = 0.0;
                    for (int q = 0; q < 6; ++q) s += Dm[r][q] * Bm[b][q][c];
                    DBm[r][c] = s;
                }
            double* kk = lK + ((size_t)b * eNoN + a) * dd;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    double BmDBm = 0.0;
                    for (int q = 0; q < 6; ++q) BmDBm += Bm[a][q][r] * DBm[q][c];
                    kk[r * dof + c] += w * ((r == c ? T1 : 0.0) + afu * BmDBm);
                }
        }
}

/* CONSTRUCT_dSOLID (svFSI/STRUCT.f:39-162): GNN at every Gauss point. */
void orc_construct_struct(int nNo, int nEl, int eNoN, const int* IEN,
                          const double* x, const double* Ag, const double* Yg,
                          const double* Dg, int tDof, const orc_eq* eq,
                          int rule, const int* rowPtr, const int* colPtr,
                          double* Val, double* R)
{
    (void)nNo;
    const int dof = 3;
    int nG;
    double wg[15], xi[15][3];
    if (eNoN == 4) { orc_rule_tet4(wg, xi); nG = 4; }
    else if (rule == 1) { orc_rule_tet10_4pt(wg, xi); nG = 4; }
    else { orc_rule_tet10(wg, xi); nG = 15; }
    double Ng[15][10], Nxig[15][10][3];
    for (int g = 0; g < nG; ++g) {
        if (eNoN == 4) orc_shape_tet4(xi[g], Ng[g], Nxig[g]);
        else orc_shape_tet10(xi[g], Ng[g], Nxig[g]);
    }
    double xl[10][3], al[30], yl[30], dl[30], Nx[10][3], ks[3][3], Jac;
    double* lR = malloc(sizeof(double) * 30);
    double* lK = malloc(sizeof(double) * 900);
    int ptr[10];
    for (int e = 0; e < nEl; ++e) {
        for (int a = 0; a < eNoN; ++a) {
            int Ac = IEN[(size_t)e * eNoN + a];
            ptr[a] = Ac;
            for (int i = 0; i < 3; ++i) {
                xl[a][i] = x[(size_t)Ac * 3 + i];
                al[a * 3 + i] = Ag[(size_t)Ac * tDof + i];
                yl[a * 3 + i] = Yg[(size_t)Ac * tDof + i];
                dl[a * 3 + i] = Dg[(size_t)Ac * tDof + i];
            }
        }
        memset(lR, 0, sizeof(double) * eNoN * dof);
        memset(lK, 0, sizeof(double) * eNoN * eNoN * 9);
        for (int g = 0; g < nG; ++g) {
            orc_gnn(eNoN, (const double (*)[3])Nxig[g], (const double (*)[3])xl, Nx, &Jac, ks);
            double w = wg[g] * Jac;
            orc_struct3d(eNoN, w, Ng[g], (const double (*)[3])Nx, al, yl, dl, 3, eq, lR, lK);
        }
        orc_doassem(eNoN, ptr, dof, lK, lR, rowPtr, colPtr, Val, R);
    }
    free(lR); free(lK);
}

/* ===================================================================== */
/* svFSILS                                                                */
/* ===================================================================== */

/* FSILS_SPARMULVV (svFSILS/SPARMUL.f:51-133), one rank.  The sums
 * associate as the reference's: for dof 1-4 (SPARMUL.f:66-113) each
 * product is added to the running row sum in turn,
 * KU = ((KU + K1 U1) + K2 U2) + ..., and for other dofs
 * (SPARMUL.f:114-126) a block row's SUM is formed first and then added. */
void orc_spmv(int nNo, int dof, const int* rowPtr, const int* colPtr,
              const double* K, const double* U, double* KU)
{
    int dd = dof * dof;
    ORC_PAR
    for (int i = 0; i < nNo; ++i) {
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j) {
            const double* k = K + (size_t)j * dd;
            const double* u = U + (size_t)colPtr[j] * dof;
            for (int l = 0; l < dof; ++l) {
                if (dof <= 4) {
                    for (int m = 0; m < dof; ++m) acc[l] += k[l * dof + m] * u[m];
                } else {
                    double s = 0.0;
                    for (int m = 0; m < dof; ++m) s += k[l * dof + m] * u[m];
                    acc[l] += s;
                }
            }
        }
        for (int l = 0; l < dof; ++l) KU[(size_t)i * dof + l] = acc[l];
    }
}

/* Transpose product K^T U of the block-CSR matrix (no reference
 * counterpart: the adjoint solve).  Block (i,j) contributes
 * K_ij^T u_i to row j. */
void orc_spmv_t(int nNo, int dof, const int* rowPtr, const int* colPtr,
                const double* K, const double* U, double* KtU)
{
    int dd = dof * dof;
    memset(KtU, 0, sizeof(double) * (size_t)nNo * dof);
    for (int i = 0; i < nNo; ++i)
        for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j) {
            const double* k = K + (size_t)j * dd;
            const double* u = U + (size_t)i * dof;
            double* y = KtU + (size_t)colPtr[j] * dof;
            for (int m = 0; m < dof; ++m) {
                double s = 0.0;
                for (int l = 0; l < dof; ++l) s += k[l * dof + m] * u[l];
                y[m] += s;
            }
        }
}

/* FSILS_NCDOTV / FSILS_DOTV one rank (svFSILS/DOT.f:56-191). */
double orc_dot(int n, const double* U, const double* V)
{
    double s = 0.0;
    ORC_PAR_SUM(s)
    for (int i = 0; i < n; ++i) s += U[i] * V[i];
    return s;
}

/* FSILS_NORMV one rank (svFSILS/NORM.f:55-113). */
double orc_norm(int n, const double* U) { return sqrt(orc_dot(n, U, U)); }

/* diagPtr of FSILS_LHS_CREATE (svFSILS/LHS.f:92-104). */
int orc_diagptr(int nNo, const int* rowPtr, const int* colPtr, int* diagPtr)
{
    for (int i = 0; i < nNo; ++i) {
        diagPtr[i] = -1;
        for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j)
            if (colPtr[j] == i) { diagPtr[i] = j; break; }
        if (diagPtr[i] < 0) return -1 - i;
    }
    return 0;
}

/* PRECONDDIAG + PREMUL/POSMUL (svFSILS/PRECOND.f:50-145, 372-489). */
void orc_precond_diag(int nNo, int dof, const int* rowPtr, const int* colPtr,
                      const int* diagPtr, const double* dirMask,
                      double* Val, double* R, double* W)
{
    int dd = dof * dof;
    ORC_PAR
    for (int a = 0; a < nNo; ++a)
        for (int i = 0; i < dof; ++i) {
            double d = Val[(size_t)diagPtr[a] * dd + i * dof + i];
            if (d == 0.0) d = 1.0;
            W[(size_t)a * dof + i] = 1.0 / sqrt(fabs(d));
        }
    if (dirMask)
        for (size_t q = 0; q < (size_t)nNo * dof; ++q) W[q] *= dirMask[q];
    ORC_PAR
    for (int a = 0; a < nNo; ++a)
        for (int j = rowPtr[a]; j < rowPtr[a + 1]; ++j)
            for (int i = 0; i < dof; ++i)
                for (int m = 0; m < dof; ++m)
                    Val[(size_t)j * dd + i * dof + m] *= W[(size_t)a * dof + i];
    for (size_t q = 0; q < (size_t)nNo * dof; ++q) R[q] *= W[q];
    ORC_PAR
    for (int a = 0; a < nNo; ++a)
        for (int j = rowPtr[a]; j < rowPtr[a + 1]; ++j) {
            int c = colPtr[j];
            for (int i = 0; i < dof; ++i)
                for (int m = 0; m < dof; ++m)
                    Val[(size_t)j * dd + i * dof + m] *= W[(size_t)c * dof + m];
        }
}

/* PRECONDRCS (svFSILS/PRECOND.f:150-368), one rank.  Dirichlet rows and
 * columns are removed with the mask (PREMUL/POSMUL, 179-197) and their
 * diagonal entries set to one (200-236).  Then up to maxiter = 10 sweeps of
 * row/column max-norm scaling (239-352): Wr, Wc = the largest |entry| of
 * each scalar row / column; the stop test max|1 - Wr| < tol and
 * max|1 - Wc| < tol with tol = 2 is evaluated before the sweep's own
 * scaling, which every sweep applies (the loop ends after it);
 * Val <- diag(1/sqrt Wr) Val diag(1/sqrt Wc), W1 *= 1/sqrt Wr,
 * W2 *= 1/sqrt Wc.  Finally R <- W1 R (355).  The solution of the scaled
 * system is unscaled with W2 (SOLVE.f:134). */
void orc_precond_rcs(int nNo, int dof, const int* rowPtr, const int* colPtr, const int* diagPtr,
                     const double* dirMask, double* Val, double* R, double* W1, double* W2)
{
    const int dd = dof * dof;
    const size_t n = (size_t)nNo * dof;
    double* Wr = malloc(sizeof(double) * n);
    double* Wc = malloc(sizeof(double) * n);
    for (size_t q = 0; q < n; ++q) {
        W1[q] = 1.0;
        W2[q] = 1.0;
        /* the face-value product renormalised to 0 / 1 (PRECOND.f:191-193) */
        Wr[q] = dirMask ? (dirMask[q] > 0.5 ? 1.0 : 0.0) : 1.0;
    }
    /* kill the Dirichlet rows and columns (PREMUL, then POSMUL) */
    for (int a = 0; a < nNo; ++a)
        for (int k = rowPtr[a]; k < rowPtr[a + 1]; ++k)
            for (int i = 0; i < dof; ++i)
                for (int j = 0; j < dof; ++j) Val[(size_t)k * dd + i * dof + j] *= Wr[(size_t)a * dof + i];
    for (size_t q = 0; q < n; ++q) R[q] = Wr[q] * R[q];
    for (int a = 0; a < nNo; ++a)
        for (int k = rowPtr[a]; k < rowPtr[a + 1]; ++k)
            for (int i = 0; i < dof; ++i)
                for (int j = 0; j < dof; ++j)
                    Val[(size_t)k * dd + i * dof + j] *= Wr[(size_t)colPtr[k] * dof + j];
    for (int a = 0; a < nNo; ++a)
        for (int i = 0; i < dof; ++i) {
            double* v = Val + (size_t)diagPtr[a] * dd + i * dof + i;
            *v = Wr[(size_t)a * dof + i] * (*v - 1.0) + 1.0;
        }
    int iter = 0, flag = 1;
    while (flag) {
        for (size_t q = 0; q < n; ++q) Wr[q] = Wc[q] = 0.0;
        ++iter;
        if (iter >= 10) flag = 0;
        for (int a = 0; a < nNo; ++a)
            for (int k = rowPtr[a]; k < rowPtr[a + 1]; ++k)
                for (int i = 0; i < dof; ++i)
                    for (int j = 0; j < dof; ++j) {
                        const double v = fabs(Val[(size_t)k * dd + i * dof + j]);
                        if (v > Wr[(size_t)a * dof + i]) Wr[(size_t)a * dof + i] = v;
                        if (v > Wc[(size_t)colPtr[k] * dof + j]) Wc[(size_t)colPtr[k] * dof + j] = v;
                    }
        double mr = 0.0, mc = 0.0;
        for (size_t q = 0; q < n; ++q) {
            mr = fmax(mr, fabs(1.0 - Wr[q]));
            mc = fmax(mc, fabs(1.0 - Wc[q]));
        }
        if (mr < 2.0 && mc < 2.0) flag = 0;
        for (size_t q = 0; q < n; ++q) {
            Wr[q] = 1.0 / sqrt(Wr[q]);
            Wc[q] = 1.0 / sqrt(Wc[q]);
        }
        for (int a = 0; a < nNo; ++a)
            for (int k = rowPtr[a]; k < rowPtr[a + 1]; ++k)
                for (int i = 0; i < dof; ++i)
                    for (int j = 0; j < dof; ++j) Val[(size_t)k * dd + i * dof + j] *= Wr[(size_t)a * dof + i];
        for (int a = 0; a < nNo; ++a)
            for (int k = rowPtr[a]; k < rowPtr[a + 1]; ++k)
                for (int i = 0; i < dof; ++i)
                    for (int j = 0; j < dof; ++j)
                        Val[(size_t)k * dd + i * dof + j] *= Wc[(size_t)colPtr[k] * dof + j];
        for (size_t q = 0; q < n; ++q) {
            W1[q] = W1[q] * Wr[q];
            W2[q] = W2[q] * Wc[q];
        }
    }
    for (size_t q = 0; q < n; ++q) R[q] = W1[q] * R[q];
    free(Wr);
    free(Wc);
}

static double wall_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* CGRADV (svFSILS/CGRAD.f:184-242). R in: RHS, out: solution. */
void orc_cgradv(int nNo, int dof, const int* rowPtr, const int* colPtr,
                const double* K, double* R, orc_subls* ls)
{
    size_t n = (size_t)nNo * dof;
    double* P = malloc(sizeof(double) * n);
    double* KP = malloc(sizeof(double) * n);
    double* X = calloc(n, sizeof(double));
    ls->callD = wall_s();
    ls->suc = 0;
    ls->iNorm = orc_norm((int)n, R);
    double eps = fmax(ls->absTol, ls->relTol * ls->iNorm);
    eps = eps * eps;
    double errO = ls->iNorm * ls->iNorm, err = errO;
    memcpy(P, R, sizeof(double) * n);
    int i;
    for (i = 1; i <= ls->mItr; ++i) {
        if (err < eps) { ls->suc = 1; break; }
        errO = err;
        orc_spmv(nNo, dof, rowPtr, colPtr, K, P, KP);
        double alpha = errO / orc_dot((int)n, P, KP);
        ORC_PAR
        for (size_t q = 0; q < n; ++q) X[q] += alpha * P[q];
        ORC_PAR
        for (size_t q = 0; q < n; ++q) R[q] += -alpha * KP[q];
        err = orc_norm((int)n, R);
        err = err * err;
        double r = errO / err;
        ORC_PAR
        for (size_t q = 0; q < n; ++q) P[q] += r * R[q];
        r = err / errO;
        ORC_PAR
        for (size_t q = 0; q < n; ++q) P[q] *= r;
    }
    memcpy(R, X, sizeof(double) * n);
    ls->itr = i - 1;
    ls->fNorm = sqrt(err);
    ls->callD = wall_s() - ls->callD;
    ls->dB = (errO < EPS) ? 0.0 : 5.0 * log(err / errO);
    free(P); free(KP); free(X);
}

/* GMRESV (svFSILS/GMRES.f:273-431) without coupled faces: restarted
 * GMRES(sD), classical Gram-Schmidt with all dots first, Pythagorean
 * sub-diagonal, Givens rotations, back substitution after each cycle. */
/* Coupled Neumann faces of one solve (FSILS_SOLVE with res(faIn) != 0 on a
 * BC_TYPE_Neu face, SOLVE.f:78-86): face f holds nodes glob[off[f] ..
 * off[f+1]-1] and the preconditioned face vector vm (valM, PRECOND.f:133-140)
 * with stride dof (fdof[f] components used), the coefficient res[f] and the
 * squared norm nS[f] of BCPRE (GMRES.f:393-429, NSSOLVER.f:307-338). */
typedef struct {
    int nC, dof;
    const int* off;
    const int* glob;
    const int* fdof;
    const double* vm;
    const double* res;
    const double* nS;
} orc_cpl;

/* ADDBCMUL (svFSILS/ADDBCMUL.f:53-114), one rank: for each coupled face in
 * order, S = coef (v . X) over its nodes and min(face dof, vd) components,
 * then Y += v S.  op 0 = BCOP_TYPE_ADD (coef = res), 1 = BCOP_TYPE_PRE
 * (coef = -res / (1 + res nS)); X and Y have vd components per node. */
static void orc_addbcmul(const orc_cpl* c, int op, int vd, const double* X, double* Y)
{
    for (int f = 0; f < c->nC; ++f) {
        const int nsd = c->fdof[f] < vd ? c->fdof[f] : vd;
        const double coef = op == 0 ? c->res[f] : -c->res[f] / (1.0 + c->res[f] * c->nS[f]);
        double S = 0.0;
        for (int a = c->off[f]; a < c->off[f + 1]; ++a)
            for (int i = 0; i < nsd; ++i) S = S + c->vm[(size_t)a * c->dof + i] * X[(size_t)c->glob[a] * vd + i];
        S = coef * S;
        for (int a = c->off[f]; a < c->off[f + 1]; ++a)
            for (int i = 0; i < nsd; ++i)
                Y[(size_t)c->glob[a] * vd + i] = Y[(size_t)c->glob[a] * vd + i] + c->vm[(size_t)a * c->dof + i] * S;
    }
}

/* BICGSV (svFSILS/BICGS.f:115-180): unpreconditioned BiCGStab on the
 * scaled system; R in: right-hand side, out: solution. */
void orc_bicgsv(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* K, double* R,
                orc_subls* ls)
{
    const size_t n = (size_t)nNo * dof;
    double* P = malloc(sizeof(double) * n);
    double* Rh = malloc(sizeof(double) * n);
    double* X = calloc(n, sizeof(double));
    double* V = malloc(sizeof(double) * n);
    double* S = malloc(sizeof(double) * n);
    double* T = malloc(sizeof(double) * n);
    ls->callD = wall_s();
    ls->suc = 0;
    double err = orc_norm((int)n, R);
    double errO = err;
    ls->iNorm = err;
    const double eps = fmax(ls->absTol, ls->relTol * err);
    double rho = err * err;
    memcpy(P, R, sizeof(double) * n);
    memcpy(Rh, R, sizeof(double) * n);
    int i;
    for (i = 1; i <= ls->mItr; ++i) {
        if (err < eps) {
            ls->suc = 1;
            break;
        }
        orc_spmv(nNo, dof, rowPtr, colPtr, K, P, V);
        const double alpha = rho / orc_dot((int)n, Rh, V);
        ORC_PAR
        for (size_t q = 0; q < n; ++q) S[q] = R[q] - alpha * V[q];
        orc_spmv(nNo, dof, rowPtr, colPtr, K, S, T);
        double omega = orc_norm((int)n, T);
        omega = orc_dot((int)n, T, S) / (omega * omega);
        ORC_PAR
        for (size_t q = 0; q < n; ++q) {
            X[q] = X[q] + alpha * P[q] + omega * S[q];
            R[q] = S[q] - omega * T[q];
        }
        errO = err;
        err = orc_norm((int)n, R);
        const double rhoO = rho;
        rho = orc_dot((int)n, R, Rh);
        const double beta = rho * alpha / (rhoO * omega);
        ORC_PAR
        for (size_t q = 0; q < n; ++q) P[q] = R[q] + beta * (P[q] - omega * V[q]);
    }
    memcpy(R, X, sizeof(double) * n);
    ls->itr = i - 1;
    ls->fNorm = err;
    ls->callD = wall_s() - ls->callD;
    ls->dB = (errO < EPS) ? 0.0 : 10.0 * log(err / errO);
    free(P); free(Rh); free(X); free(V); free(S); free(T);
}

static void gmresv_impl(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* K, double* R,
                        orc_subls* ls, const orc_cpl* cp);

void orc_gmresv(int nNo, int dof, const int* rowPtr, const int* colPtr,
                const double* K, double* R, orc_subls* ls)
{
    gmresv_impl(nNo, dof, rowPtr, colPtr, K, R, ls, NULL);
}

/* GMRESV with the coupled faces' ADDBCMUL(ADD) after every product
 * (GMRES.f:317-331; its BCOP_TYPE_PRE branch is off, flag = .FALSE.) */
static void gmresv_impl(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* K, double* R,
                        orc_subls* ls, const orc_cpl* cp)
{
    size_t n = (size_t)nNo * dof;
    int sD = ls->sD;
    double* u = malloc(sizeof(double) * n * (sD + 1));
    double* X = calloc(n, sizeof(double));
    double* h = calloc((size_t)(sD + 1) * sD, sizeof(double)); /* h[(i)*(sD+1)+j] = h(j,i) */
    double* y = malloc(sizeof(double) * sD);
    double* c = malloc(sizeof(double) * sD);
    double* s = malloc(sizeof(double) * sD);
    double* err = malloc(sizeof(double) * (sD + 1));
#define H(j, i) h[(size_t)(i) * (sD + 1) + (j)]
    ls->callD = wall_s();
    ls->suc = 0;
    double eps = orc_norm((int)n, R);
    ls->iNorm = eps;
    ls->fNorm = eps;
    eps = fmax(ls->absTol, ls->relTol * eps);
    ls->itr = 0;
    if (ls->iNorm <= ls->absTol) {
        /* GMRES.f:307-311 returns before R = X: R keeps the scaled RHS */
        ls->callD = EPS;
        ls->dB = 0.0;
        goto done;
    }
    for (int l = 1; l <= ls->mItr; ++l) {
        ls->dB = ls->fNorm;
        ls->itr++;
        orc_spmv(nNo, dof, rowPtr, colPtr, K, X, u);
        if (cp) orc_addbcmul(cp, 0, dof, X, u);
        ORC_PAR
        for (size_t q = 0; q < n; ++q) u[q] = R[q] - u[q];
        err[0] = orc_norm((int)n, u);
        ORC_PAR
        for (size_t q = 0; q < n; ++q) u[q] = u[q] / err[0];
        int i;
        for (i = 1; i <= sD; ++i) {
            ls->itr++;
            double* ui = u + n * (i - 1);
            double* un = u + n * i;
            orc_spmv(nNo, dof, rowPtr, colPtr, K, ui, un);
            if (cp) orc_addbcmul(cp, 0, dof, ui, un);
            for (int j = 1; j <= i + 1; ++j) H(j - 1, i - 1) = orc_dot((int)n, u + n * (j - 1), un);
            for (int j = 1; j <= i; ++j) {
                double hj = H(j - 1, i - 1);
                const double* uj = u + n * (j - 1);
                ORC_PAR
                for (size_t q = 0; q < n; ++q) un[q] += -hj * uj[q];
                H(i, i - 1) = H(i, i - 1) - hj * hj;
            }
            H(i, i - 1) = sqrt(fabs(H(i, i - 1)));
            double rr = 1.0 / H(i, i - 1);
            ORC_PAR
            for (size_t q = 0; q < n; ++q) un[q] *= rr;
            for (int j = 1; j <= i - 1; ++j) {
                double tmp = c[j - 1] * H(j - 1, i - 1) + s[j - 1] * H(j, i - 1);
                H(j, i - 1) = -s[j - 1] * H(j - 1, i - 1) + c[j - 1] * H(j, i - 1);
                H(j - 1, i - 1) = tmp;
            }
            double tmp = sqrt(H(i - 1, i - 1) * H(i - 1, i - 1) + H(i, i - 1) * H(i, i - 1));
            c[i - 1] = H(i - 1, i - 1) / tmp;
            s[i - 1] = H(i, i - 1) / tmp;
            H(i - 1, i - 1) = tmp;
            H(i, i - 1) = 0.0;
            err[i] = -s[i - 1] * err[i - 1];
            err[i - 1] = c[i - 1] * err[i - 1];
            if (fabs(err[i]) < eps) { ls->suc = 1; break; }
        }
        if (i > sD) i = sD;
        for (int j = 0; j < i; ++j) y[j] = err[j];
        for (int j = i; j >= 1; --j) {
            for (int k = j + 1; k <= i; ++k) y[j - 1] = y[j - 1] - H(j - 1, k - 1) * y[k - 1];
            y[j - 1] = y[j - 1] / H(j - 1, j - 1);
        }
        for (int j = 1; j <= i; ++j) {
            const double* uj = u + n * (j - 1);
            ORC_PAR
            for (size_t q = 0; q < n; ++q) X[q] += y[j - 1] * uj[q];
        }
        ls->fNorm = fabs(err[i]);
        if (ls->suc) break;
    }
    memcpy(R, X, sizeof(double) * n);
    ls->callD = wall_s() - ls->callD;
    ls->dB = 10.0 * log(ls->fNorm / ls->dB);
done:
#undef H
    free(u); free(X); free(h); free(y); free(c); free(s); free(err);
}

/* ===================================================================== */
/* NSSOLVER (svFSILS/NSSOLVER.f:52-383) and its parts                     */
/* ===================================================================== */

/* GMRES (svFSILS/GMRES.f:51-166): restarted GMRES(sD) for K X = R from
 * X = 0.  Unlike GMRESV the first cycle starts from R itself (no product),
 * R is left unchanged, and ls->itr accumulates over calls. */
static void orc_gmres_x(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* K,
                        const double* R, double* X, orc_subls* ls, const orc_cpl* cp)
{
    size_t n = (size_t)nNo * dof;
    int sD = ls->sD;
    double* u = malloc(sizeof(double) * n * (sD + 1));
    double* h = calloc((size_t)(sD + 1) * sD, sizeof(double));
    double* y = malloc(sizeof(double) * sD);
    double* c = malloc(sizeof(double) * sD);
    double* s = malloc(sizeof(double) * sD);
    double* err = malloc(sizeof(double) * (sD + 1));
    double* unc = cp ? malloc(sizeof(double) * n) : NULL;
#define H(j, i) h[(size_t)(i) * (sD + 1) + (j)]
    double time = wall_s();
    ls->suc = 0;
    double eps = 0.0;
    memset(X, 0, sizeof(double) * n);
    for (int l = 1; l <= ls->mItr; ++l) {
        if (l == 1) {
            memcpy(u, R, sizeof(double) * n);
        } else {
            ls->itr++;
            orc_spmv(nNo, dof, rowPtr, colPtr, K, X, u);
            if (cp) orc_addbcmul(cp, 0, dof, X, u);
            ORC_PAR
            for (size_t q = 0; q < n; ++q) u[q] = R[q] - u[q];
        }
        if (cp) {
            memcpy(unc, u, sizeof(double) * n);
            orc_addbcmul(cp, 1, dof, unc, u);
        }
        err[0] = orc_norm((int)n, u);
        if (l == 1) {
            eps = err[0];
            if (eps <= ls->absTol) {
                ls->callD = EPS;
                ls->dB = 0.0;
                goto done;
            }
            ls->iNorm = eps;
            ls->fNorm = eps;
            eps = fmax(ls->absTol, ls->relTol * eps);
        }
        ls->dB = ls->fNorm;
        ORC_PAR
        for (size_t q = 0; q < n; ++q) u[q] = u[q] / err[0];
        int i;
        for (i = 1; i <= sD; ++i) {
            ls->itr++;
            double* ui = u + n * (i - 1);
            double* un = u + n * i;
            orc_spmv(nNo, dof, rowPtr, colPtr, K, ui, un);
            if (cp) {
                orc_addbcmul(cp, 0, dof, ui, un);
                memcpy(unc, un, sizeof(double) * n);
                orc_addbcmul(cp, 1, dof, unc, un);
            }
            for (int j = 1; j <= i + 1; ++j) H(j - 1, i - 1) = orc_dot((int)n, u + n * (j - 1), un);
            for (int j = 1; j <= i; ++j) {
                double hj = H(j - 1, i - 1);
                const double* uj = u + n * (j - 1);
                ORC_PAR
                for (size_t q = 0; q < n; ++q) un[q] += -hj * uj[q];
                H(i, i - 1) = H(i, i - 1) - hj * hj;
            }
            H(i, i - 1) = sqrt(fabs(H(i, i - 1)));
            double rr = 1.0 / H(i, i - 1);
            ORC_PAR
            for (size_t q = 0; q < n; ++q) un[q] *= rr;
            for (int j = 1; j <= i - 1; ++j) {
                double tmp = c[j - 1] * H(j - 1, i - 1) + s[j - 1] * H(j, i - 1);
                H(j, i - 1) = -s[j - 1] * H(j - 1, i - 1) + c[j - 1] * H(j, i - 1);
                H(j - 1, i - 1) = tmp;
            }
            double tmp = sqrt(H(i - 1, i - 1) * H(i - 1, i - 1) + H(i, i - 1) * H(i, i - 1));
            c[i - 1] = H(i - 1, i - 1) / tmp;
            s[i - 1] = H(i, i - 1) / tmp;
            H(i - 1, i - 1) = tmp;
            H(i, i - 1) = 0.0;
            err[i] = -s[i - 1] * err[i - 1];
            err[i - 1] = c[i - 1] * err[i - 1];
            if (fabs(err[i]) < eps) { ls->suc = 1; break; }
        }
        if (i > sD) i = sD;
        for (int j = 0; j < i; ++j) y[j] = err[j];
        for (int j = i; j >= 1; --j) {
            for (int k = j + 1; k <= i; ++k) y[j - 1] = y[j - 1] - H(j - 1, k - 1) * y[k - 1];
            y[j - 1] = y[j - 1] / H(j - 1, j - 1);
        }
        for (int j = 1; j <= i; ++j) {
            const double* uj = u + n * (j - 1);
            ORC_PAR
            for (size_t q = 0; q < n; ++q) X[q] = X[q] + y[j - 1] * uj[q];
        }
        ls->fNorm = fabs(err[i]);
        if (ls->suc) break;
    }
    ls->callD = wall_s() - time + ls->callD;
    ls->dB = 10.0 * log(ls->fNorm / ls->dB);
done:
#undef H
    free(u); free(h); free(y); free(c); free(s); free(err); free(unc);
}

/* FSILS_SPARMULVS / SV / SS (svFSILS/SPARMUL.f:135-297), one rank:
 * vector-to-scalar (K(nsd,nnz), U(nsd,nNo) -> KU(nNo)), scalar-to-vector
 * and scalar-to-scalar block products, in the reference's summation
 * order. */
static void orc_spmv_vs(int nNo, int nsd, const int* rowPtr, const int* colPtr, const double* K,
                        const double* U, double* KU)
{
    ORC_PAR
    for (int i = 0; i < nNo; ++i) {
        double a = 0.0;
        for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j) {
            const double* k = K + (size_t)j * nsd;
            const double* u = U + (size_t)colPtr[j] * nsd;
            double t = a + k[0] * u[0];
            for (int m = 1; m < nsd; ++m) t = t + k[m] * u[m];
            a = t;
        }
        KU[i] = a;
    }
}

static void orc_spmv_sv(int nNo, int nsd, const int* rowPtr, const int* colPtr, const double* K,
                        const double* U, double* KU)
{
    ORC_PAR
    for (int i = 0; i < nNo; ++i) {
        double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j)
            for (int m = 0; m < nsd; ++m) a[m] = a[m] + K[(size_t)j * nsd + m] * U[colPtr[j]];
        for (int m = 0; m < nsd; ++m) KU[(size_t)i * nsd + m] = a[m];
    }
}

static void orc_spmv_ss(int nNo, const int* rowPtr, const int* colPtr, const double* K, const double* U,
                        double* KU)
{
    ORC_PAR
    for (int i = 0; i < nNo; ++i) {
        double a = 0.0;
        for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j) a = a + K[j] * U[colPtr[j]];
        KU[i] = a;
    }
}

/* CGRAD_SCHUR (svFSILS/CGRAD.f:51-123): CG on the Schur complement
 * S P = L P - D (G P) (D = Gt = -G^T blockwise, so S = L + G^T G); R in:
 * right-hand side, out: solution; ls->itr accumulates. */
static void orc_cgrad_schur(int nNo, int nsd, const int* rowPtr, const int* colPtr, const double* D,
                            const double* G, const double* L, double* R, orc_subls* ls, const orc_cpl* cp)
{
    double* unc = cp ? malloc(sizeof(double) * (size_t)nNo * nsd) : NULL;
    double* X = calloc(nNo, sizeof(double));
    double* P = malloc(sizeof(double) * nNo);
    double* SP = malloc(sizeof(double) * nNo);
    double* DGP = malloc(sizeof(double) * nNo);
    double* GP = malloc(sizeof(double) * (size_t)nNo * nsd);
    double time = wall_s();
    ls->suc = 0;
    ls->iNorm = orc_norm(nNo, R);
    double eps = pow(fmax(ls->absTol, ls->relTol * ls->iNorm), 2.0);
    double errO = ls->iNorm * ls->iNorm, err = errO;
    memcpy(P, R, sizeof(double) * nNo);
    int i;
    for (i = 1; i <= ls->mItr; ++i) {
        if (err < eps) { ls->suc = 1; break; }
        errO = err;
        orc_spmv_sv(nNo, nsd, rowPtr, colPtr, G, P, GP);
        if (cp) {
            memcpy(unc, GP, sizeof(double) * (size_t)nNo * nsd);
            orc_addbcmul(cp, 1, nsd, unc, GP);
        }
        orc_spmv_vs(nNo, nsd, rowPtr, colPtr, D, GP, DGP);
        orc_spmv_ss(nNo, rowPtr, colPtr, L, P, SP);
        ORC_PAR
        for (int q = 0; q < nNo; ++q) SP[q] = SP[q] + (-1.0) * DGP[q];
        double alpha = errO / orc_dot(nNo, P, SP);
        ORC_PAR
        for (int q = 0; q < nNo; ++q) X[q] = X[q] + alpha * P[q];
        ORC_PAR
        for (int q = 0; q < nNo; ++q) R[q] = R[q] + (-alpha) * SP[q];
        err = orc_norm(nNo, R);
        err = err * err;
        ORC_PAR
        for (int q = 0; q < nNo; ++q) P[q] = P[q] + (errO / err) * R[q];
        ORC_PAR
        for (int q = 0; q < nNo; ++q) P[q] = P[q] * (err / errO);
    }
    memcpy(R, X, sizeof(double) * nNo);
    ls->fNorm = sqrt(err);
    ls->callD = wall_s() - time + ls->callD;
    ls->itr = ls->itr + i - 1;
    ls->dB = (errO < EPS) ? 0.0 : 5.0 * log(err / errO);
    free(X); free(P); free(SP); free(DGP); free(GP); free(unc);
}

/* GE (svFSILS/GE.f:51-151): A x = B for the leading N x N block of A
 * (column-major, leading dimension nV) after symmetric diagonal scaling;
 * partial pivoting; closed forms for N = 1, 2.  B is overwritten with x;
 * returns 0 (and B = 0) for a singular system. */
static int orc_ge(int nV, int N, const double* A, double* B)
{
#define AA(i, j) A[(size_t)(j) * nV + (i)]
    if (N <= 0) return 0;
    double* W = malloc(sizeof(double) * N);
    for (int i = 0; i < N; ++i) {
        if (fabs(AA(i, i)) < DBL_MIN) {
            for (int q = 0; q < N; ++q) B[q] = 0.0;
            free(W);
            return 0;
        }
        W[i] = 1.0 / sqrt(fabs(AA(i, i)));
    }
    double* C = malloc(sizeof(double) * N * (N + 1)); /* C(i,j) at i*(N+1)+j */
#define CC(i, j) C[(size_t)(i) * (N + 1) + (j)]
    for (int i = 0; i < N; ++i) {
        for (int j = 0; j < N; ++j) CC(i, j) = W[i] * W[j] * AA(i, j);
        CC(i, N) = W[i] * B[i];
    }
    int ok = 1;
    if (N == 1) {
        B[0] = CC(0, 1) / CC(0, 0);
        B[0] = B[0] * W[0];
    } else if (N == 2) {
        double pivot = CC(0, 0) * CC(1, 1) - CC(1, 0) * CC(0, 1);
        if (fabs(pivot) < DBL_EPSILON) {
            B[0] = B[1] = 0.0;
            ok = 0;
        } else {
            B[0] = (CC(0, 2) * CC(1, 1) - CC(1, 2) * CC(0, 1)) / pivot;
            B[1] = (CC(1, 2) * CC(0, 0) - CC(0, 2) * CC(1, 0)) / pivot;
            B[0] = W[0] * B[0];
            B[1] = W[1] * B[1];
        }
    } else {
        for (int m = 0; m < N - 1 && ok; ++m) {
            int ipv = m;
            double pivot = fabs(CC(m, m));
            for (int i = m + 1; i < N; ++i)
                if (fabs(CC(i, m)) > pivot) { ipv = i; pivot = fabs(CC(i, m)); }
            if (fabs(pivot) < DBL_EPSILON) {
                for (int q = 0; q < N; ++q) B[q] = 0.0;
                ok = 0;
                break;
            }
            if (ipv != m)
                for (int j = m; j <= N; ++j) {
                    double t = CC(m, j);
                    CC(m, j) = CC(ipv, j);
                    CC(ipv, j) = t;
                }
            for (int i = m + 1; i < N; ++i) {
                double f = CC(i, m) / CC(m, m);
                CC(i, m) = 0.0;
                for (int j = m + 1; j <= N; ++j) CC(i, j) = CC(i, j) - f * CC(m, j);
            }
        }
        if (ok) {
            for (int j = N - 1; j >= 0; --j) {
                for (int i = j + 1; i < N; ++i) CC(j, N) = CC(j, N) - CC(j, i) * CC(i, N);
                CC(j, N) = CC(j, N) / CC(j, j);
            }
            for (int i = 0; i < N; ++i) B[i] = W[i] * CC(i, N);
        }
    }
#undef CC
#undef AA
    free(C);
    free(W);
    return ok;
}

/* GE exported for its test */
int orc_ge_test(int nV, int N, const double* A, double* B) { return orc_ge(nV, N, A, B); }

/* NSSOLVER (svFSILS/NSSOLVER.f:52-383), one rank, nsd = dof - 1 = 3,
 * no coupled faces.  DEPART (236-303) splits each 4x4 block into the
 * momentum block mK (3x3), the gradient column mG, the divergence row mD
 * and mL, and Gt(:,trans(k)) = -mG(:,k).  Each outer iteration i adds two
 * directions: [0; P_i] with P_i = S^-1 (Rc - D K^-1 Rm) (CGRAD_SCHUR) and
 * [U_i; 0] with U_i = K^-1 (Rm - G P_i) (GMRES), their images MU/MP under
 * the full matrix, the Gram system A xB = B over all directions so far
 * (GE), and the residual.  ls holds RI, GM and CG in FSILS_lsType order. */
static int nssolver_impl(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* Val, double* Ri,
                         orc_ls* ls, const orc_cpl* cp);

int orc_nssolver(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* Val, double* Ri,
                 orc_ls* ls)
{
    return nssolver_impl(nNo, dof, rowPtr, colPtr, Val, Ri, ls, NULL);
}

/* NSSOLVER with the coupled faces: ADDBCMUL(ADD) on K U (NSSOLVER.f:128),
 * ADD and PRE inside GMRES, PRE inside CGRAD_SCHUR */
static int nssolver_impl(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* Val, double* Ri,
                         orc_ls* ls, const orc_cpl* cp)
{
    if (dof != 4) return -3;
    const int nsd = 3;
    const size_t n = (size_t)nNo, nnz = (size_t)rowPtr[nNo];
    const int iBmax = ls->RI.mItr, nB = 2 * iBmax;
    double* mK = malloc(sizeof(double) * nnz * 9);
    double* mG = malloc(sizeof(double) * nnz * 3);
    double* mD = malloc(sizeof(double) * nnz * 3);
    double* mL = malloc(sizeof(double) * nnz);
    double* Gt = malloc(sizeof(double) * nnz * 3);
    ORC_PAR
    for (size_t k = 0; k < nnz; ++k) {
        const double* t = Val + k * 16;
        const int kk[9] = {0, 1, 2, 4, 5, 6, 8, 9, 10};
        for (int m = 0; m < 9; ++m) mK[k * 9 + m] = t[kk[m]];
        mG[k * 3 + 0] = t[3];
        mG[k * 3 + 1] = t[7];
        mG[k * 3 + 2] = t[11];
        mD[k * 3 + 0] = t[12];
        mD[k * 3 + 1] = t[13];
        mD[k * 3 + 2] = t[14];
        mL[k] = t[15];
    }
    ORC_PAR
    for (int i = 0; i < nNo; ++i)
        for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j) {
            const int k = colPtr[j];
            for (int l = rowPtr[k]; l < rowPtr[k + 1]; ++l)
                if (colPtr[l] == i) {
                    for (int m = 0; m < 3; ++m) Gt[(size_t)l * 3 + m] = -mG[(size_t)j * 3 + m];
                    break;
                }
        }
    double* Rm = malloc(sizeof(double) * n * nsd);
    double* Rc = malloc(sizeof(double) * n);
    double* Rmi = malloc(sizeof(double) * n * nsd);
    double* Rci = malloc(sizeof(double) * n);
    double* U = malloc(sizeof(double) * n * nsd * iBmax);
    double* P = malloc(sizeof(double) * n * iBmax);
    double* MU = malloc(sizeof(double) * n * nsd * nB);
    double* MP = malloc(sizeof(double) * n * nB);
    double* A = calloc((size_t)nB * nB, sizeof(double));
    double* B = calloc(nB, sizeof(double));
    double* xB = calloc(nB, sizeof(double));
    double* oldxB = calloc(nB, sizeof(double));
    ORC_PAR
    for (size_t a = 0; a < n; ++a) {
        for (int m = 0; m < nsd; ++m) Rmi[a * nsd + m] = Ri[a * dof + m];
        Rci[a] = Ri[a * dof + nsd];
    }
    memcpy(Rm, Rmi, sizeof(double) * n * nsd);
    memcpy(Rc, Rci, sizeof(double) * n);
    double eps = sqrt(pow(orc_norm((int)(n * nsd), Rm), 2.0) + pow(orc_norm((int)n, Rc), 2.0));
    ls->RI.iNorm = eps;
    ls->RI.fNorm = eps * eps;
    ls->CG.callD = 0.0;
    ls->GM.callD = 0.0;
    ls->CG.itr = 0;
    ls->GM.itr = 0;
    ls->RI.callD = wall_s();
    ls->RI.suc = 0;
    eps = fmax(ls->RI.absTol, ls->RI.relTol * eps);
#define VU(i) (U + (size_t)((i) - 1) * n * nsd)
#define VP(i) (P + (size_t)((i) - 1) * n)
#define VMU(i) (MU + (size_t)((i) - 1) * n * nsd)
#define VMP(i) (MP + (size_t)((i) - 1) * n)
#define AM(j, k) A[(size_t)((k) - 1) * nB + ((j) - 1)]
    int i, iB = 0, iBB = 0;
    for (i = 1; i <= ls->RI.mItr; ++i) {
        iB = 2 * i - 1;
        iBB = 2 * i;
        ls->RI.dB = ls->RI.fNorm;
        orc_gmres_x(nNo, nsd, rowPtr, colPtr, mK, Rm, VU(i), &ls->GM, cp);
        orc_spmv_vs(nNo, nsd, rowPtr, colPtr, mD, VU(i), VP(i));
        ORC_PAR
        for (size_t q = 0; q < n; ++q) VP(i)[q] = Rc[q] - VP(i)[q];
        orc_cgrad_schur(nNo, nsd, rowPtr, colPtr, Gt, mG, mL, VP(i), &ls->CG, cp);
        orc_spmv_sv(nNo, nsd, rowPtr, colPtr, mG, VP(i), VMU(iB));
        ORC_PAR
        for (size_t q = 0; q < n * nsd; ++q) VMU(iBB)[q] = Rm[q] - VMU(iB)[q];
        orc_gmres_x(nNo, nsd, rowPtr, colPtr, mK, VMU(iBB), VU(i), &ls->GM, cp);
        orc_spmv(nNo, nsd, rowPtr, colPtr, mK, VU(i), VMU(iBB));
        if (cp) orc_addbcmul(cp, 0, nsd, VU(i), VMU(iBB));
        orc_spmv_ss(nNo, rowPtr, colPtr, mL, VP(i), VMP(iB));
        orc_spmv_vs(nNo, nsd, rowPtr, colPtr, mD, VU(i), VMP(iBB));
        for (int k = iB; k <= iBB; ++k) {
            for (int j = 1; j <= k; ++j) {
                const double t = orc_dot((int)(n * nsd), VMU(j), VMU(k)) + orc_dot((int)n, VMP(j), VMP(k));
                AM(j, k) = t;
                AM(k, j) = t;
            }
            B[k - 1] = orc_dot((int)(n * nsd), VMU(k), Rmi) + orc_dot((int)n, VMP(k), Rci);
        }
        memcpy(xB, B, sizeof(double) * nB);
        if (orc_ge(nB, iBB, A, xB)) {
            memcpy(oldxB, xB, sizeof(double) * nB);
        } else {
            memcpy(xB, oldxB, sizeof(double) * nB);
            if (i > 1) {
                iB -= 2;
                iBB -= 2;
            }
            break;
        }
        double sxb = 0.0;
        for (int k = 0; k < iBB; ++k) sxb += xB[k] * B[k];
        ls->RI.fNorm = ls->RI.iNorm * ls->RI.iNorm - sxb;
        if (ls->RI.fNorm < eps * eps) {
            ls->RI.suc = 1;
            break;
        }
        ORC_PAR
        for (size_t q = 0; q < n * nsd; ++q) Rm[q] = Rmi[q] - xB[0] * VMU(1)[q];
        ORC_PAR
        for (size_t q = 0; q < n; ++q) Rc[q] = Rci[q] - xB[0] * VMP(1)[q];
        for (int j = 2; j <= iBB; ++j) {
            for (size_t q = 0; q < n * nsd; ++q) Rm[q] = Rm[q] - xB[j - 1] * VMU(j)[q];
            ORC_PAR
            for (size_t q = 0; q < n; ++q) Rc[q] = Rc[q] - xB[j - 1] * VMP(j)[q];
        }
    }
    if (i > ls->RI.mItr) {
        ls->RI.itr = ls->RI.mItr;
    } else {
        ls->RI.itr = i;
        ORC_PAR
        for (size_t q = 0; q < n; ++q) Rc[q] = Rci[q] - xB[0] * VMP(1)[q];
        for (int j = 2; j <= iBB; ++j)
            for (size_t q = 0; q < n; ++q) Rc[q] = Rc[q] - xB[j - 1] * VMP(j)[q];
    }
    ls->Resc = (int)lround(100.0 * pow(orc_norm((int)n, Rc), 2.0) / ls->RI.fNorm);
    ls->Resm = 100 - ls->Resc;
    ORC_PAR
    for (size_t q = 0; q < n * nsd; ++q) Rmi[q] = xB[1] * VU(1)[q];
    ORC_PAR
    for (size_t q = 0; q < n; ++q) Rci[q] = xB[0] * VP(1)[q];
    for (int k = 2; k <= ls->RI.itr; ++k) {
        const int b1 = 2 * k - 1, b2 = 2 * k;
        ORC_PAR
        for (size_t q = 0; q < n * nsd; ++q) Rmi[q] = Rmi[q] + xB[b2 - 1] * VU(k)[q];
        ORC_PAR
        for (size_t q = 0; q < n; ++q) Rci[q] = Rci[q] + xB[b1 - 1] * VP(k)[q];
    }
    ls->RI.callD = wall_s() - ls->RI.callD;
    ls->RI.dB = 5.0 * log(ls->RI.fNorm / ls->RI.dB);
    if (ls->Resc < 0 || ls->Resm < 0) {
        ls->Resc = 0;
        ls->Resm = 0;
        ls->RI.dB = 0.0;
        ls->RI.fNorm = 0.0;
    }
    ls->RI.fNorm = sqrt(ls->RI.fNorm);
    ORC_PAR
    for (size_t a = 0; a < n; ++a) {
        for (int m = 0; m < nsd; ++m) Ri[a * dof + m] = Rmi[a * nsd + m];
        Ri[a * dof + nsd] = Rci[a];
    }
#undef VU
#undef VP
#undef VMU
#undef VMP
#undef AM
    free(mK); free(mG); free(mD); free(mL); free(Gt);
    free(Rm); free(Rc); free(Rmi); free(Rci); free(U); free(P); free(MU); free(MP);
    free(A); free(B); free(xB); free(oldxB);
    return 0;
}

/* FSILS_SOLVE with the full solver record (RI, GM, CG): LS_type 796 (NS),
 * 797 (GMRES) or 798 (CG); prec 701 or 709. */
int orc_fsils_solve_ls(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* dirMask,
                       int LS_type, int prec, orc_ls* ls, double* Ri, double* Val)
{
    if (LS_type != 796)
        return orc_fsils_solve_prec(nNo, dof, rowPtr, colPtr, dirMask, LS_type, prec, &ls->RI, Ri, Val);
    int* diagPtr = malloc(sizeof(int) * nNo);
    if (orc_diagptr(nNo, rowPtr, colPtr, diagPtr) != 0) { free(diagPtr); return -1; }
    size_t n = (size_t)nNo * dof;
    double* W = malloc(sizeof(double) * n);
    double* W1 = malloc(sizeof(double) * n);
    if (prec == 709)
        orc_precond_rcs(nNo, dof, rowPtr, colPtr, diagPtr, dirMask, Val, Ri, W1, W);
    else
        orc_precond_diag(nNo, dof, rowPtr, colPtr, diagPtr, dirMask, Val, Ri, W);
    int rc = orc_nssolver(nNo, dof, rowPtr, colPtr, Val, Ri, ls);
    for (size_t q = 0; q < n; ++q) Ri[q] = W[q] * Ri[q];
    free(W); free(W1); free(diagPtr);
    return rc;
}

int orc_fsils_solve_cpl(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* dirMask,
                        int LS_type, int prec, orc_ls* ls, int nC, const int* cOff, const int* cGlob,
                        const int* cDof, const double* cVal, const double* cRes, double* Ri, double* Val)
{
    if (nC <= 0) return orc_fsils_solve_ls(nNo, dof, rowPtr, colPtr, dirMask, LS_type, prec, ls, Ri, Val);
    /* PRECONDRCS leaves valM unset (PRECOND.f:356-364 are commented out) */
    if (prec != 701) return -4;
    int* diagPtr = malloc(sizeof(int) * nNo);
    if (orc_diagptr(nNo, rowPtr, colPtr, diagPtr) != 0) { free(diagPtr); return -1; }
    const size_t n = (size_t)nNo * dof, T = (size_t)cOff[nC];
    double* W = malloc(sizeof(double) * n);
    double* vm = calloc(T * dof, sizeof(double));
    double* nS = calloc(nC, sizeof(double));
    orc_precond_diag(nNo, dof, rowPtr, colPtr, diagPtr, dirMask, Val, Ri, W);
    /* valM = val W (PRECOND.f:133-140); nS over dof - 1 components (BCPRE) */
    for (int f = 0; f < nC; ++f) {
        const int nv = cDof[f] < dof ? cDof[f] : dof;
        const int ns = cDof[f] < dof - 1 ? cDof[f] : dof - 1;
        for (int a = cOff[f]; a < cOff[f + 1]; ++a) {
            for (int i = 0; i < nv; ++i)
                vm[(size_t)a * dof + i] = cVal[(size_t)a * dof + i] * W[(size_t)cGlob[a] * dof + i];
            for (int i = 0; i < ns; ++i) nS[f] = nS[f] + vm[(size_t)a * dof + i] * vm[(size_t)a * dof + i];
        }
    }
    orc_cpl c = {nC, dof, cOff, cGlob, cDof, vm, cRes, nS};
    int rc = 0;
    if (LS_type == 796) rc = nssolver_impl(nNo, dof, rowPtr, colPtr, Val, Ri, ls, &c);
    else if (LS_type == 797) gmresv_impl(nNo, dof, rowPtr, colPtr, Val, Ri, &ls->RI, &c);
    else if (LS_type == 798) orc_cgradv(nNo, dof, rowPtr, colPtr, Val, Ri, &ls->RI); /* CGRADV has no coupled terms */
    else if (LS_type == 795) orc_bicgsv(nNo, dof, rowPtr, colPtr, Val, Ri, &ls->RI); /* nor BICGSV */
    else rc = -2;
    for (size_t q = 0; q < n; ++q) Ri[q] = W[q] * Ri[q];
    free(W); free(vm); free(nS); free(diagPtr);
    return rc;
}

/* FSILS_SOLVE (svFSILS/SOLVE.f:51-143), one rank, PRECOND_FSILS, no
 * coupled Neumann faces: scaling, CG or GMRES, un-scaling x = W y. */
int orc_fsils_solve_prec(int nNo, int dof, const int* rowPtr, const int* colPtr,
                         const double* dirMask, int LS_type, int prec, orc_subls* ls,
                         double* Ri, double* Val)
{
    int* diagPtr = malloc(sizeof(int) * nNo);
    if (orc_diagptr(nNo, rowPtr, colPtr, diagPtr) != 0) { free(diagPtr); return -1; }
    size_t n = (size_t)nNo * dof;
    double* W = malloc(sizeof(double) * n);
    double* W1 = malloc(sizeof(double) * n);
    if (prec == 709) /* PRECOND_RCS: unscale with W2 */
        orc_precond_rcs(nNo, dof, rowPtr, colPtr, diagPtr, dirMask, Val, Ri, W1, W);
    else
        orc_precond_diag(nNo, dof, rowPtr, colPtr, diagPtr, dirMask, Val, Ri, W);
    if (LS_type == 797) orc_gmresv(nNo, dof, rowPtr, colPtr, Val, Ri, ls);
    else if (LS_type == 798) orc_cgradv(nNo, dof, rowPtr, colPtr, Val, Ri, ls);
    else if (LS_type == 795) orc_bicgsv(nNo, dof, rowPtr, colPtr, Val, Ri, ls);
    else { free(W); free(W1); free(diagPtr); return -2; }
    for (size_t q = 0; q < n; ++q) Ri[q] = W[q] * Ri[q];
    free(W); free(W1); free(diagPtr);
    return 0;
}

int orc_fsils_solve(int nNo, int dof, const int* rowPtr, const int* colPtr,
                    const double* dirMask, int LS_type, orc_subls* ls,
                    double* Ri, double* Val)
{
    return orc_fsils_solve_prec(nNo, dof, rowPtr, colPtr, dirMask, LS_type, 701, ls, Ri, Val);
}

/* ===================================================================== */
/* Adjoint helper                                                         */
/* ===================================================================== */

/* g_e = -lam_e^T (dK_e/dE_e) u_e for the LELAS tet4 tangent.  The
 * stiffness part of LELAS3D is linear in E (lambda, mu proportional to E)
 * and the mass part amd*N_a*N_b/mu*wl = w*am*rho*N_a*N_b does not depend
 * on E, so dK_e/dE_e = (K_e(E_e) - M_e)/E_e = K_e(1) - M_e/... computed
 * here as the stiffness tangent at unit modulus. */
void orc_lelas_dkde_dot(int nEl, const int* IEN, const double* x,
                        const orc_eq* eq, const double* E_el,
                        const double* lam, const double* u, double* g)
{
    (void)E_el;
    const int eNoN = 4;
    double wg[4], xi[4][3], Ng[4][4], Nxi[4][3];
    orc_rule_tet4(wg, xi);
    for (int gq = 0; gq < 4; ++gq) orc_shape_tet4(xi[gq], Ng[gq], Nxi);
    orc_eq eq0 = *eq;
    eq0.rho = 0.0; /* drop the mass part: dK/dE is the unit-modulus stiffness */
    double xl[4][3], Nx[4][3], ks[3][3], Jac = 0.0, lR[12], lK[144], zero[12] = {0};
    for (int e = 0; e < nEl; ++e) {
        for (int a = 0; a < eNoN; ++a)
            for (int i = 0; i < 3; ++i) xl[a][i] = x[(size_t)IEN[e * 4 + a] * 3 + i];
        memset(lK, 0, sizeof(lK));
        memset(lR, 0, sizeof(lR));
        orc_gnn(eNoN, (const double (*)[3])Nxi, (const double (*)[3])xl, Nx, &Jac, ks);
        for (int gq = 0; gq < 4; ++gq)
            orc_lelas3d(eNoN, wg[gq] * Jac, Ng[gq], (const double (*)[3])Nx, zero, zero, 3, &eq0, 1.0, lR, lK);
        double s = 0.0;
        for (int b = 0; b < 4; ++b)
            for (int a = 0; a < 4; ++a) {
                const double* k = lK + (b * 4 + a) * 9;
                const double* la = lam + (size_t)IEN[e * 4 + a] * 3;
                const double* ub = u + (size_t)IEN[e * 4 + b] * 3;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) s += la[i] * k[i * 3 + j] * ub[j];
            }
        g[e] = -s;
    }
}
