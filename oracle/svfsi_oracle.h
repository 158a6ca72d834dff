/*
 * svfsi_oracle.h -- CPU restatement of the svFSI / svFSILS per-Newton-step
 * linear-system path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path (svfsi_amd/libsvfsi_amd.so): only tests/, the smoke()
 * entry point and the CPU-baseline leg of bench.py may load it.  The
 * product path never calls it and has no CPU fallback.
 *
 * Written from the mathematics of the reference routines (cited per
 * function, paths relative to Code/Source/ of SimVascular/svFSI), FP64
 * and int32 throughout like the reference (svFSI/TYPEMOD.f:56-58,
 * svFSILS/FSILS_TYPEDEF.h:55-57).
 *
 * Layout conventions (0-based C indices, Fortran column-major memory):
 *   x[nNo][3]          == x(nsd,tnNo)
 *   IEN[nEl][eNoN]     == IEN(eNoN,nEl), node ids 0-based
 *   state[nNo][tDof]   == Ag/Yg/Dg(tDof,tnNo)
 *   Val[nnz][dof*dof]  == Val(dof*dof,nnz); block row-major: (i,j) at i*dof+j
 *   R[nNo][dof]        == R(dof,tnNo)
 *   lK[b][a][dof*dof]  == lK(dof*dof,a,b)
 *   rowPtr[nNo+1], colPtr[nnz] 0-based, columns sorted ascending per row
 *
 * Parity pinning of each function is listed in the test files
 * tests/test_oracle_*.py; functions whose result no closed form, finite
 * difference or brute-force check fixes are marked "parity unpinned" below.
 */
#ifndef SVFSI_ORACLE_H
#define SVFSI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Equation / material parameters (COMMOD eq(cEq) and dmn props). */
typedef struct {
    double am, af, beta, gam, dt;   /* generalized-alpha (INITIALIZE.f:76-139) */
    double rho;                     /* density (solid or fluid) */
    double E, nu;                   /* LELAS: elasticity modulus, Poisson ratio */
    double fb[3];                   /* body force f_x, f_y, f_z */
    double mu;                      /* FLUID: constant viscosity (viscType_Const) */
    double bstab;                   /* FLUID: backflow stabilization coefficient */
    double C10, Kpen;               /* STRUCT nHK: C10 = mu/2, penalty (bulk modulus) */
    int volType;                    /* STRUCT: 1 Quad, 2 ST91, 3 M94 (GETSVOLP) */
    int pad;
    /* FLUID viscosity model (GETVISCOSITY, FLUID.f:1690-1739): 698 (or 0)
     * constant mu; 697 Carreau-Yasuda (mu_i = mu, mu_o, lam, a, n); 696
     * Casson (mu_i = mu, mu_o, lam) */
    int viscType, pad2;
    double mu_o, lam, visc_a, visc_n;
} orc_eq;

/* FSILS_subLsType (FSILS_STRUCT.h:151-180) */
typedef struct {
    int suc, mItr, sD, itr, cM, cN, cD, reserve;
    double absTol, relTol, iNorm, fNorm, dB, callD;
} orc_subls;

/* FSILS_lsType (FSILS_STRUCT.h:182-204): the NS solver's outer (RI) and
 * inner (GM, CG) records */
typedef struct {
    int foC, LS_type, Resm, Resc;
    orc_subls GM, CG, RI;
} orc_ls;

/* ---------------- NN.f: quadrature, shape functions, GNN ---------------- */
void orc_rule_tet4(double w[4], double xi[4][3]);
void orc_rule_tet10(double w[15], double xi[15][3]);
void orc_rule_tet10_4pt(double w[4], double xi[4][3]);
void orc_rule_tri3(double w[3], double xi[3][2]);
void orc_shape_tet4(const double xi[3], double N[4], double Nxi[4][3]);
void orc_shape_tet10(const double xi[3], double N[10], double Nxi[10][3]);
void orc_shape_tri3(const double xi[2], double N[3], double Nxi[3][2]);
void orc_gnn(int eNoN, const double (*Nxi)[3], const double (*xl)[3],
             double (*Nx)[3], double* Jac, double ks[3][3]);
int  orc_iszero(double a);

/* ---------------- LHSA.f: sparsity pattern, DOASSEM ---------------- */
int  orc_lhsa_count(int nNo, int nEl, int eNoN, const int* IEN, int* rowPtr);
void orc_lhsa_fill(int nNo, int nEl, int eNoN, const int* IEN,
                   const int* rowPtr, int* colPtr);
void orc_doassem(int d, const int* eqN, int dof, const double* lK,
                 const double* lR, const int* rowPtr, const int* colPtr,
                 double* Val, double* R);

/* ---------------- element kernels ---------------- */
void orc_lelas3d(int eNoN, double w, const double* N, const double (*Nx)[3],
                 const double* al, const double* dl, int tDof,
                 const orc_eq* eq, double E, double* lR, double* lK);
void orc_fluid3d_m(int vms, int eNoN, int eNoNq, double w, const double ks[3][3], const double* N,
                   const double* Nq, const double (*Nx)[3], const double (*Nqx)[3], const double (*Nxx)[6],
                   const double* al, const double* yl, int tDof, const orc_eq* eq, double mu, double* lR,
                   double* lK);
void orc_fluid3d_c(int vms, int eNoN, int eNoNq, double w, const double ks[3][3], const double* N,
                   const double* Nq, const double (*Nx)[3], const double (*Nqx)[3], const double (*Nxx)[6],
                   const double* al, const double* yl, int tDof, const orc_eq* eq, double mu, double* lR,
                   double* lK);
void orc_getpk2cc_nhk(const orc_eq* eq, const double F[3][3],
                      double S[3][3], double Dm[6][6]);
void orc_struct3d(int eNoN, double w, const double* N, const double (*Nx)[3],
                  const double* al, const double* yl, const double* dl,
                  int tDof, const orc_eq* eq, double* lR, double* lK);

/* GETVISCOSITY (svFSI/FLUID.f:1690-1739): mu, mu_s and d mu / d gamma at
 * shear rate gamma_in (Casson returns the raised rate in *gamma_out) */
void orc_getviscosity(const orc_eq* eq, double mu_i, double gamma_in, double* gamma_out, double* mu,
                      double* mu_s, double* mu_x);

/* ---------------- construct loops (element loop + DOASSEM) ---------------- */
/* rule: 0 = reference rule for the element (tet4: 4-pt, tet10: 15-pt),
 *       1 = tet10 4-point rule (the configuration's cheaper rule).
 * prop_el: per-element E (LELAS) or mu (FLUID), NULL = uniform eq value. */
void orc_construct_lelas(int nNo, int nEl, const int* IEN, const double* x,
                         const double* Ag, const double* Dg, int tDof,
                         const orc_eq* eq, const double* E_el,
                         const int* rowPtr, const int* colPtr,
                         double* Val, double* R);
void orc_construct_fluid(int nNo, int nEl, const int* IEN, const double* x,
                         const double* Ag, const double* Yg, int tDof,
                         const orc_eq* eq, const double* mu_el,
                         const int* rowPtr, const int* colPtr,
                         double* Val, double* R);
/* CONSTRUCT_FLUID with the Taylor-Hood basis (nFs = 2): tet10 velocity,
 * tet4 pressure, no VMS terms; IEN[nEl][10].  Returns nonzero if GNNxx
 * meets a singular element.  THOOD_ValRC (FS.f:352-407) is separate, as
 * svFSI calls it after the assembly. */
int orc_construct_fluid_th(int nNo, int nEl, const int* IEN, const double* x,
                           const double* Ag, const double* Yg, int tDof,
                           const orc_eq* eq, const double* mu_el,
                           const int* rowPtr, const int* colPtr,
                           double* Val, double* R);
void orc_thood_valrc(int nNo, int nEl, const int* IEN, const int* rowPtr, const int* colPtr, double* Val,
                     double* R);
/* GNNxx (svFSI/NN.f:1568-1678) and the TET10 second-derivative table. */
int orc_gnnxx(int eNoN, const double (*Nxi)[3], const double (*Nxi2)[6], const double (*xl)[3],
              const double (*Nx)[3], double (*Nxx)[6]);
void orc_shape_tet10_xx(double Nxi2[10][6]);
/* Viscosity sensitivity (no reference counterpart): g[e] = lam_e .
 * (dR_e/dmu_e - dK_e/dmu_e du_e), derivative by complex step. */
void orc_fluid_dmu_dot(int nEl, const int* IEN, const double* x, const double* Ag,
                       const double* Yg, int tDof, const orc_eq* eq, const double* mu_el,
                       const double* lam, const double* du, double* g);
void orc_construct_struct(int nNo, int nEl, int eNoN, const int* IEN,
                          const double* x, const double* Ag, const double* Yg,
                          const double* Dg, int tDof, const orc_eq* eq,
                          int rule, const int* rowPtr, const int* colPtr,
                          double* Val, double* R);
/* Face (tri3) terms: traction (SETBCTRACL) and outlet backflow (BFLUID via
 * BASSEMNEUBC with h = 0).  fIEN[nFe][3] face connectivity, fEl[nFe] the
 * adjacent volume element (for the outward orientation, GNNB). */
void orc_traction_tri3(int nFe, const int* fIEN, const int* fEl,
                       const int* IEN, int eNoN, const double* x,
                       const double h[3], int dof, const int* rowPtr,
                       const int* colPtr, double* Val, double* R);
void orc_backflow_tri3(int nFe, const int* fIEN, const int* fEl,
                       const int* IEN, int eNoN, const double* x,
                       const double* Yg, int tDof, const orc_eq* eq,
                       const int* rowPtr, const int* colPtr,
                       double* Val, double* R);

/* ---------------- svFSILS solver ---------------- */
void   orc_spmv(int nNo, int dof, const int* rowPtr, const int* colPtr,
                const double* K, const double* U, double* KU);
void   orc_spmv_t(int nNo, int dof, const int* rowPtr, const int* colPtr,
                  const double* K, const double* U, double* KtU);
double orc_dot(int n, const double* U, const double* V);
double orc_norm(int n, const double* U);
int    orc_diagptr(int nNo, const int* rowPtr, const int* colPtr, int* diagPtr);
void   orc_precond_diag(int nNo, int dof, const int* rowPtr, const int* colPtr,
                        const int* diagPtr, const double* dirMask,
                        double* Val, double* R, double* W);
void   orc_cgradv(int nNo, int dof, const int* rowPtr, const int* colPtr,
                  const double* K, double* R, orc_subls* ls);
void   orc_gmresv(int nNo, int dof, const int* rowPtr, const int* colPtr,
                  const double* K, double* R, orc_subls* ls);
void   orc_bicgsv(int nNo, int dof, const int* rowPtr, const int* colPtr,
                  const double* K, double* R, orc_subls* ls);
/* FSILS_SOLVE for LS_TYPE_GMRES (797), LS_TYPE_CG (798) or, through
 * orc_fsils_solve_prec, LS_TYPE_BICGS (795), with PRECOND_FSILS.
 * dirMask[nNo][dof]: product of the Dirichlet face values (1 elsewhere). */
int    orc_fsils_solve(int nNo, int dof, const int* rowPtr, const int* colPtr,
                       const double* dirMask, int LS_type, orc_subls* ls,
                       double* Ri, double* Val);

/* ---------------- adjoint (no counterpart in the reference) ------------- */
/* dK_e/dE_e . u contributions: g_e = -lambda_e^T (dK_e/dE_e) u_e, with
 * dK/dE = K_stiff/E (the mass part does not depend on E). */
void orc_lelas_dkde_dot(int nEl, const int* IEN, const double* x,
                        const orc_eq* eq, const double* E_el,
                        const double* lam, const double* u, double* g);

/* NSSOLVER (svFSILS/NSSOLVER.f:52-383) on the scaled system, dof 4; and
 * FSILS_SOLVE with the full record (LS_type 796 NS, 797 GMRES, 798 CG) */
int orc_nssolver(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* Val, double* Ri,
                 orc_ls* ls);
int orc_fsils_solve_ls(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* dirMask,
                       int LS_type, int prec, orc_ls* ls, double* Ri, double* Val);
/* FSILS_SOLVE with coupled Neumann faces (res(faIn) != 0, SOLVE.f:78-86)
 * and PRECOND_FSILS: face f = nodes cGlob[cOff[f] .. cOff[f+1]-1] (0-based),
 * face vector cVal (stride dof, cDof[f] components), coefficient cRes[f].
 * ADDBCMUL enters GMRESV and NSSOLVER; CGRADV ignores the faces, as the
 * reference does.  Returns -4 for PRECOND_RCS (valM unset there). */
int orc_fsils_solve_cpl(int nNo, int dof, const int* rowPtr, const int* colPtr, const double* dirMask,
                        int LS_type, int prec, orc_ls* ls, int nC, const int* cOff, const int* cGlob,
                        const int* cDof, const double* cVal, const double* cRes, double* Ri, double* Val);
/* GE (svFSILS/GE.f:51-151), A column-major with leading dimension nV */
int orc_ge_test(int nV, int N, const double* A, double* B);

/* threads of the solver kernels (CPU baseline timing only; default 1) */
void orc_set_threads(int n);

/* PRECONDRCS (svFSILS/PRECOND.f:150-368) and FSILS_SOLVE with either
 * preconditioner (701 PRECOND_FSILS, 709 PRECOND_RCS) */
void orc_precond_rcs(int nNo, int dof, const int* rowPtr, const int* colPtr, const int* diagPtr,
                     const double* dirMask, double* Val, double* R, double* W1, double* W2);
int orc_fsils_solve_prec(int nNo, int dof, const int* rowPtr, const int* colPtr,
                         const double* dirMask, int LS_type, int prec, orc_subls* ls,
                         double* Ri, double* Val);

#ifdef __cplusplus
}
#endif

#endif
