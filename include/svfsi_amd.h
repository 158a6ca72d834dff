/*
 * svfsi_amd.h -- C ABI of libsvfsi_amd.so, the MI355X (gfx950) implementation
 * of svFSI's per-Newton-step linear-system path: element tangent/residual
 * construction + DOASSEM scatter into svFSILS block-CSR, then the svFSILS
 * Krylov solve with diagonal (Jacobi) scaling.
 *
 * Reference interfaces replaced (paths under Code/Source/ of SimVascular/svFSI):
 *   svFSILS/FSILS_api_c.h:52-64      C prototypes fsils_*_ (gfortran mangling)
 *   svFSILS/FSILS_API.h:51-135       Fortran explicit interfaces FSILS_*
 *   svFSILS/FSILS_struct_c.h:115-140 C layout of FSILS_lsType / subLsType
 *   svFSI/EQASSEM.f:38-87            GLOBALEQASSEM element loop (no C ABI in
 *                                    the reference; precedent: the Trilinos
 *                                    sink trilinos_doassem_,
 *                                    svFSI/trilinos_linear_solver.h:201-208)
 *
 * Conventions
 *   - Plain C types only; no torch or HIP types in signatures.  A stream is
 *     passed as void* (a hipStream_t, NULL = the default stream).
 *   - Functions with a _dev suffix take DEVICE pointers (memory resident in
 *     HBM, e.g. from hipMalloc or a torch tensor); the others take HOST
 *     pointers and copy.
 *   - Indices that cross the ABI follow the reference: 1-based rowPtr /
 *     colPtr / gNodes / IEN, Fortran column-major arrays: Ri(dof,nNo),
 *     Val(dof*dof,nnz) with each dof x dof block row-major
 *     (Val((i-1)*dof+j,k) = dR_i/dU_j), x(3,nNo), IEN(eNoN,nEl).
 *   - Every entry point returns 0 on success or a negative SVFSI_ERR_* code;
 *     svfsi_last_error() describes the failure.  The reference STOPs on
 *     misuse (e.g. svFSILS/LS.f:60-63, SOLVE.f:78-81); this library returns
 *     the error instead.  Non-convergence is not an error: as in the
 *     reference it is reported through ls->RI.suc = 0.
 *   - There is no CPU fallback: every compute entry point runs HIP kernels
 *     and fails with SVFSI_ERR_NODEVICE when no GPU is present.
 */
#ifndef SVFSI_AMD_H
#define SVFSI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Codes (svFSILS/FSILS_STRUCT.h:53-62)                                    */
/* ---------------------------------------------------------------------- */
#define LS_TYPE_CG     798
#define LS_TYPE_GMRES  797
#define LS_TYPE_NS     796
#define LS_TYPE_BICGS  795
#define PRECOND_FSILS  701
#define PRECOND_RCS    709
#define BC_TYPE_Dir    0
#define BC_TYPE_Neu    1

#define SVFSI_OK            0
#define SVFSI_ERR_ARG      -1  /* bad argument (reference: FATAL ERROR STOP) */
#define SVFSI_ERR_HIP      -2  /* HIP runtime error */
#define SVFSI_ERR_STATE    -3  /* object not created / already created */
#define SVFSI_ERR_NOTSUP   -4  /* combination the reference supports but this build does not */
#define SVFSI_ERR_NODEVICE -5  /* no HIP device */

/* MPI_Aint of the reference's C structs and prototypes: taken from mpi.h
 * when the caller includes it first, otherwise the pointer-sized signed
 * integer MPI_Aint is in MPICH and Open MPI. */
#if defined(FSILS_C_H) || defined(MPI_VERSION)
typedef MPI_Aint fsils_aint_t;
#else
typedef intptr_t fsils_aint_t;
#endif

/* When the reference's own FSILS_c.h has been included first (FSILS_C_H
 * defined), its struct types are used, and the fsils_*_ prototypes further
 * down redeclare its functions: a C compiler then checks them against the
 * reference's. */
#ifndef FSILS_C_H

/* FSILS_commuType, same layout as svFSILS/FSILS_struct_c.h:60-68 */
typedef struct {
    int foC;            /* created (USE) */
    int masF;           /* this rank is the master (USE) */
    int master;         /* master rank (USE) */
    int task;           /* rank (USE) */
    int tF;             /* rank + 1 (USE) */
    int nTasks;         /* number of ranks (USE) */
    fsils_aint_t comm;  /* MPI communicator (IN) */
} FSILS_commuType;

/* FSILS_subLsType, same layout as svFSILS/FSILS_struct_c.h:115-130 */
typedef struct {
    int suc;       /* successful solve (OUT) */
    int mItr;      /* max iterations / restarts (IN) */
    int sD;        /* Krylov space dimension (IN) */
    int itr;       /* iterations performed (OUT) */
    int cM, cN, cD;/* counters (OUT) */
    int reserve;
    double absTol; /* absolute tolerance (IN) */
    double relTol; /* relative tolerance (IN) */
    double iNorm;  /* initial residual norm (OUT) */
    double fNorm;  /* final residual norm (OUT) */
    double dB;     /* residual reduction (OUT) */
    double callD;  /* wall time of the call, s (OUT) */
} FSILS_subLsType;

/* FSILS_lsType, same layout as svFSILS/FSILS_struct_c.h:132-140 */
typedef struct {
    int foC;
    int LS_type;
    int Resm, Resc;
    FSILS_subLsType GM;
    FSILS_subLsType CG;
    FSILS_subLsType RI;
} FSILS_lsType;

#endif /* FSILS_C_H */

const char* svfsi_last_error(void);
/* Number of visible HIP devices (0 when none). */
int svfsi_device_count(void);

/* ---------------------------------------------------------------------- */
/* svFSILS drop-in layer                                                   */
/* ---------------------------------------------------------------------- */

/* Replaces fsils_ls_create_ (FSILS_api_c.h:58; FSILS_LS_CREATE, svFSILS/LS.f:50-119).
 * Optional arguments are passed as NULL.  Fills defaults exactly as the
 * reference: GMRES relTol 0.1, mItr 4, sD 250; CG relTol 1e-2, mItr 1000;
 * absTol 1e-10. */
int fsils_gpu_ls_create(FSILS_lsType* ls, int LS_type, const double* relTol,
                        const double* absTol, const int* maxItr,
                        const int* dimKry, const double* relTolIn,
                        const double* absTolIn, const int* maxItrIn);

/* Replaces fsils_lhs_create_c_ (FSILS_api_c.h:55; FSILS_LHS_CREATE,
 * svFSILS/LHS.f:51-111 one-rank branch).  rowPtr(nNo+1), colPtr(nnz) are
 * 1-based host arrays; the pattern is copied to the device, diagPtr is
 * built, and a transpose-position map is built for the adjoint.  gNodes is
 * accepted for signature parity (one rank: identity map, LHS.f:106). */
int fsils_gpu_lhs_create(void** lhs, int gnNo, int nNo, int nnz,
                         const int* gNodes, const int* rowPtr,
                         const int* colPtr, int nFaces);
/* Replaces fsils_lhs_free_ (FSILS_api_c.h:56; LHS.f). */
int fsils_gpu_lhs_free(void* lhs);

/* Replaces fsils_bc_create_ (FSILS_api_c.h:61; FSILS_BC_CREATE,
 * svFSILS/BC.f:50-121).  faIn is 1-based; gNodes 1-based; Val(dof,nNo)
 * may be NULL (then 0, a homogeneous Dirichlet mask).  For a BC_TYPE_Neu
 * face Val is the face vector (the integral of N n, svFSI/BAFINI.f:514-533)
 * that couples through res(faIn) in the solve. */
int fsils_gpu_bc_create(void* lhs, int faIn, int nNo, int dof, int BC_type,
                        const int* gNodes, const double* Val);
/* Replaces fsils_bc_free_ (FSILS_api_c.h:62). */
int fsils_gpu_bc_free(void* lhs, int faIn);

/* Replaces fsils_solve_ (FSILS_api_c.h:64; FSILS_SOLVE, svFSILS/SOLVE.f:51-143)
 * with HOST arrays: Ri(dof,nNo) is the RHS on entry and the solution on
 * exit; Val(dof*dof,nnz) is left scaled in place (W K W), as the reference
 * does (SOLVE.f:57, PRECOND.f:124-130).  prec PRECOND_FSILS (diagonal
 * scaling) or PRECOND_RCS (row/column equilibration, PRECOND.f:150-368);
 * LS_type GMRES, CG or BICGS (BICGSV, svFSILS/BICGS.f:115-180) (dof >= 2),
 * or NS (dof 4: NSSOLVER,
 * svFSILS/NSSOLVER.f:52-383, with the ls%GM and ls%CG records, filling
 * ls%Resm/Resc).  incL(nFaces) may be NULL (all faces included).
 * res(nFaces): a BC_TYPE_Neu face with res != 0 adds res v v^T (v its face
 * vector scaled by W) to the GMRES and NS operators (ADDBCMUL,
 * svFSILS/ADDBCMUL.f:53-114); CG ignores it, as CGRADV does.  res may be
 * NULL only when no BC_TYPE_Neu face exists (SOLVE.f:74-77). */
int fsils_gpu_solve(void* lhs, FSILS_lsType* ls, int dof, double* Ri,
                    double* Val, int prec, const int* incL, const double* res);

/* Same as fsils_gpu_solve with DEVICE arrays dRi, dVal (no host copies). */
int fsils_gpu_solve_dev(void* lhs, FSILS_lsType* ls, int dof, double* dRi,
                        double* dVal, int prec, const int* incL,
                        const double* res, void* stream);

/* Mixed-precision form of fsils_gpu_solve_dev (no reference counterpart;
 * the reference is FP64-only, svFSILS/FSILS_TYPEDEF.h:55-57).  Same
 * arguments, scaling, convergence test and outputs.  The solver is always
 * flexible GMRES in FP64 (outer residual f - Ku, classical Gram-Schmidt
 * Arnoldi in one pass with the explicit norm (two passes on saddle-point
 * systems), Hessenberg/Givens least squares, solution), with
 * restart length ls->RI.sD (30 when 0), whose right preconditioner works on
 * an FP32 copy of the scaled matrix: a fixed GMRES polynomial of the
 * (block-Jacobi scaled) operator, one V-cycle of an aggregation multigrid
 * (dof 4 with an FP32 matrix above 128 MiB), or one cycle of an inner FP32 GMRES(30)
 * stopped at a relative reduction of 0.1 (chosen, without block Jacobi and
 * stopped at 0.01, when unmasked rows of the scaled matrix have a zero
 * diagonal: a saddle-point system such as the Taylor-Hood fluid).  LS_type GMRES, CG and BICGS all
 * select this one solver (it replaces each of them; their defaults of
 * fsils_gpu_ls_create still apply); LS_TYPE_NS is refused
 * (SVFSI_ERR_NOTSUP).  The reported residual is the FP64 one in the
 * reference's scaled norm; ls->RI.mItr bounds the outer restarts.
 * ls->RI.itr counts outer (FP64) Arnoldi steps and restarts;
 * ls->RI.cM counts all SpMVs (inner FP32 plus outer FP64). */
int fsils_gpu_solve_mixed_dev(void* lhs, FSILS_lsType* ls, int dof,
                              double* dRi, double* dVal, int prec,
                              const int* incL, const double* res, void* stream);

/* No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  fsils_gpu_solve_dev restricted to LS_TYPE_CG with
 * PRECOND_FSILS (CGRADV with diagonal scaling; coupled faces play no part in
 * CG), with the same arguments, results and outputs, that calls no hipMalloc,
 * hipFree or other allocating routine.  Its work memory (W, the Dirichlet
 * mask, the CG vectors, partial sums and state) is one fixed device array of
 * svfsi_noalloc_scratch_bytes() bytes (64 MiB: about 1.6 million unknowns,
 * nNo * dof), which each GPU the library runs on holds until the process
 * ends.  Larger systems are refused with SVFSI_ERR_ARG before any GPU work.
 * All calls on one GPU share that array: calls that overlap in time on one
 * GPU (from two host threads or two streams) race and must not be made. */
int fsils_gpu_solve_cg_noalloc_dev(void* lhs, FSILS_lsType* ls, int dof, double* dRi, double* dVal,
                                   const int* incL, void* stream);
size_t svfsi_noalloc_scratch_bytes(void);

/* ---------------------------------------------------------------------- */
/* The reference's C API, svFSILS/FSILS_api_c.h:52-64, same prototypes     */
/* ---------------------------------------------------------------------- */
/* A driver written against FSILS_c.h links this library in place of
 * svFSILS.  lhs is an opaque handle (the reference's is a pointer to its
 * SAVEd Fortran lhs, LHS.f:330-348; here each fsils_lhs_create_c_ call makes
 * a new one).  Optional Fortran arguments are NULL pointers.  As in the
 * reference these return nothing and stop the program on misuse, printing
 * the reason and "FSILS: FATAL ERROR"; see svfsi_fsils_fatal below. */

/* FSILS_COMMU_CREATE (svFSILS/COMMU.f:50-86): reads the Fortran MPI handle
 * (the low 4 bytes of *commi); size and rank come from the process's MPI
 * library when one is loaded and initialised, else 1 and 0. */
void fsils_commu_create_(FSILS_commuType* commu, fsils_aint_t* commi);
/* FSILS_COMMU_FREE (COMMU.f:88-96). */
void fsils_commu_free_(FSILS_commuType* commu);
/* FSILS_LHS_CREATE_C (svFSILS/LHS.f:330-348) over fsils_gpu_lhs_create; a
 * communicator of more than one rank stops with "not supported". */
void fsils_lhs_create_c_(void** lhs, FSILS_commuType* commu, int* gnNo, int* nNo, int* nnz, int* gNodes,
                         int* rowPtr, int* colPtr, int* nFaces);
/* FSILS_LHS_FREE (LHS.f:295-328). */
void fsils_lhs_free_(void* lhs);
/* FSILS_LS_CREATE (svFSILS/LS.f:50-119) over fsils_gpu_ls_create. */
void fsils_ls_create_(FSILS_lsType* ls, int* LS_type, double* relTol, double* absTol, int* maxItr, int* dimKry,
                      double* relTolIn, double* absTolIn, int* maxItrIn);
/* FSILS_LS_FREE (LS.f:121-133). */
void fsils_ls_free_(FSILS_lsType* ls);
/* FSILS_BC_CREATE (svFSILS/BC.f:50-121) over fsils_gpu_bc_create. */
void fsils_bc_create_(void* lhs, int* faIn, int* nNo, int* dof, int* BC_type, int* gNodes, double* Val);
/* FSILS_BC_FREE (BC.f:195-214). */
void fsils_bc_free_(void* lhs, int* faIn);
/* FSILS_SOLVE (svFSILS/SOLVE.f:51-143) over fsils_gpu_solve: host arrays,
 * Ri(dof,nNo) RHS in and solution out, Val left scaled in place. */
void fsils_solve_(void* lhs, FSILS_lsType* ls, int* dof, double* Ri, double* Val, int* prec, int* incL,
                  double* res);

/* No counterpart in the reference: on = 0 makes the fsils_*_ functions
 * above return after a failure instead of stopping the program; the status
 * of the last call (0 or an SVFSI_ERR_* code) is then svfsi_fsils_status()
 * and the reason svfsi_last_error().  Returns the previous setting. */
int svfsi_fsils_fatal(int on);
int svfsi_fsils_status(void);

/* ---------------------------------------------------------------------- */
/* Device memory for callers without a HIP binding (the Fortran driver)    */
/* ---------------------------------------------------------------------- */
/* No counterpart in the reference (its arrays live in host memory).  They
 * wrap hipMalloc / hipFree / synchronous hipMemcpy on the current device, so
 * a driver can keep Ag/Yg/Dg, R and Val resident in HBM across Newton
 * iterations and call the _dev entry points. */
int svfsi_dev_malloc(void** dptr, size_t bytes);
int svfsi_dev_free(void* dptr);
int svfsi_memcpy_h2d(void* dst, const void* src, size_t bytes);
int svfsi_memcpy_d2h(void* dst, const void* src, size_t bytes);

/* ---------------------------------------------------------------------- */
/* Kernel-level entry points of the solve (device arrays)                  */
/* ---------------------------------------------------------------------- */

/* FSILS_SPARMULVV (svFSILS/SPARMUL.f:51-133): KU = K U, one rank. */
int svfsi_spmv_dev(void* lhs, int dof, const double* dK, const double* dU,
                   double* dKU, void* stream);
/* Same product with FP32 values and vectors (mixed path). */
int svfsi_spmv_f32_dev(void* lhs, int dof, const float* dK, const float* dU,
                       float* dKU, void* stream);
/* Transpose product KU = K^T U (adjoint solve; no reference counterpart). */
int svfsi_spmv_t_dev(void* lhs, int dof, const double* dK, const double* dU,
                     double* dKU, void* stream);
/* Aggregation multigrid, the inner preconditioner the mixed entry points use
 * with SVFSI_INNER_AMG=1 (no reference counterpart: svFSILS has diagonal
 * and row/column scaling only).  svfsi_amg_aggregate: the greedy aggregation
 * of a 0-based CSR graph (rowPtr nNo+1, colPtr) into aggregates of at most
 * cap nodes (a node left alone joins the smallest neighbouring one); agg
 * receives each node's aggregate; returns the number of aggregates (< 0 on
 * error).  svfsi_amg_levels: the node counts of the hierarchy the last mixed
 * solve on lhs built (finest first) into nodes[0..maxLevels); returns the
 * number of levels (0: none). */
int svfsi_amg_aggregate(int nNo, const int* rowPtr, const int* colPtr, int cap, int* agg);
int svfsi_amg_levels(void* lhs, int* nodes, int maxLevels);

/* Residual dR = dB - K dU (transpose != 0: dB - K^T dU) with every row's sum
 * carried in double-double (FMA-split products, TwoSum accumulation) and
 * rounded once: the residual of iterative refinement on ill-conditioned
 * systems (adjoint gradients; no reference counterpart).  dR may alias dB
 * but not dU. */
int svfsi_resid_dd_dev(void* lhs, int dof, const double* dK, const double* dU,
                       const double* dB, double* dR, int transpose, void* stream);
/* FSILS_DOTV / FSILS_NORMV one rank (svFSILS/DOT.f:56-113, NORM.f:55-113):
 * deterministic two-stage reduction; the result is written to *out (host). */
int svfsi_dot_dev(int n, const double* dU, const double* dV, double* out,
                  void* stream);
/* PRECONDDIAG (svFSILS/PRECOND.f:50-145): scales dVal and dR in place and
 * writes W(dof,nNo) to dW, honouring the Dirichlet faces of lhs. */
int svfsi_precond_diag_dev(void* lhs, int dof, double* dVal, double* dR,
                           double* dW, const int* incL, void* stream);
/* PRECONDRCS (svFSILS/PRECOND.f:150-368): scales dVal and dR in place
 * (Dirichlet rows/columns removed, up to 10 row/column max-norm sweeps)
 * and writes the row factors W1 and column factors W2 (dof,nNo); the
 * solution of the scaled system is unscaled with W2 (SOLVE.f:134).  Needs
 * a structurally symmetric pattern. */
int svfsi_precond_rcs_dev(void* lhs, int dof, double* dVal, double* dR,
                          double* dW1, double* dW2, const int* incL, void* stream);

/* ---------------------------------------------------------------------- */
/* Sparsity pattern                                                        */
/* ---------------------------------------------------------------------- */

/* LHSA (svFSI/LHSA.f:38-262): node-adjacency block pattern of a mesh.
 * IEN(eNoN,nEl) 1-based.  Output 1-based rowPtr(nNo+1) and colPtr(nnz),
 * columns sorted ascending and unique per row, diagonal included.  Call
 * with colPtr = NULL to obtain *nnz and rowPtr, then again with colPtr of
 * length *nnz.  Host setup routine (runs once per mesh; needs no GPU). */
int svfsi_lhsa(int nNo, int nEl, int eNoN, const int* IEN, int* nnz,
               int* rowPtr, int* colPtr);

/* ---------------------------------------------------------------------- */
/* Assembly (replaces the element loop of GLOBALEQASSEM, EQASSEM.f:45-84)  */
/* ---------------------------------------------------------------------- */

/* Equation parameters (eq(cEq) / dmn props of COMMOD, svFSI/MOD.f). */
typedef struct {
    double am, af, beta, gam, dt; /* generalized-alpha, INITIALIZE.f:76-139 */
    double rho;                   /* solid or fluid density */
    double E, nu;                 /* LELAS (LELAS.f:146-155) */
    double fb[3];                 /* body force f_x,f_y,f_z */
    double mu;                    /* FLUID constant viscosity */
    double bstab;                 /* FLUID backflow stabilization (default 0.2) */
    double C10, Kpen;             /* STRUCT nHK (MATMODELS.f:136-150) */
    int volType;                  /* 1 Quad, 2 ST91 (default), 3 M94 */
    int pad;
    /* FLUID viscosity model (GETVISCOSITY, FLUID.f:1690-1739; READVISCMODEL,
     * READFILES.f:2878-2914): viscType SVFSI_VISC_CONST (mu), SVFSI_VISC_CY
     * (Carreau-Yasuda: mu_i = mu, mu_o, lam, visc_a, visc_n) or
     * SVFSI_VISC_CASS (Casson: mu_i = mu, mu_o, lam = low shear-rate
     * threshold).  0 is read as SVFSI_VISC_CONST. */
    int viscType;
    int pad2;
    double mu_o, lam, visc_a, visc_n;
} svfsi_eq_params;

#define SVFSI_VISC_CONST 698  /* viscType_Const (CONSTS.f:201-202) */
#define SVFSI_VISC_CY    697  /* viscType_CY */
#define SVFSI_VISC_CASS  696  /* viscType_Cass */

#define SVFSI_PHYS_FLUID   1  /* CONSTRUCT_FLUID, FLUID.f:40-190 (tet4 VMS) */
#define SVFSI_PHYS_LELAS   2  /* CONSTRUCT_LELAS, LELAS.f:38-128 (tet4) */
#define SVFSI_PHYS_STRUCT  3  /* CONSTRUCT_dSOLID, STRUCT.f:39-162 (nHK) */

/* STRUCT (tet10) element kernel, ORed into svfsi_assemble_dev's quad.  No
 * counterpart in the reference (STRUCT3D is one FP64 code path there):
 *   VALU   -- FP64 vector units throughout;
 *   MFMA64 -- B^T Dm B on FP64 matrix cores (the default);
 *   MFMA32 -- B^T Dm B on FP32 matrix cores with FP32 accumulation (the
 *             rest, and the accumulation into Val, stays FP64). */
/* FLUID on a tet10 mesh: the Taylor-Hood (P2P1) basis of svFSI's nFs = 2
 * (SETTHOODFS/GETTHOODFS, svFSI/FS.f:201-346): quadratic velocity on the
 * tet10 nodes, linear pressure on the corner nodes, no VMS terms; the
 * assembly ends with THOOD_ValRC (FS.f:352-407), which sets the pressure
 * rows of the mid-edge nodes to the identity.  Passed as quad. */
#define SVFSI_FLUID_P2P1    0x400

#define SVFSI_STRUCT_AUTO   0x000
#define SVFSI_STRUCT_VALU   0x100
#define SVFSI_STRUCT_MFMA64 0x200
#define SVFSI_STRUCT_MFMA32 0x300

/* Device-resident mesh bound to an lhs pattern: coordinates, connectivity,
 * a deterministic element colouring and the per-element DOASSEM scatter map
 * (the column search of LHSA.f:282-292, done once).  IEN(eNoN,nEl) 1-based
 * host array; x(3,nNo) host array.  eNoN 4 (tet4) or 10 (tet10). */
int svfsi_mesh_create(void** mesh, void* lhs, int nNo, int nEl, int eNoN,
                      const int* IEN, const double* x);
int svfsi_mesh_free(void* mesh);
/* Number of element colours (for reporting). */
int svfsi_mesh_ncolors(void* mesh);

/* Boundary faces of the mesh (tri3): fIEN(3,nFe) 1-based node ids, fEl(nFe)
 * 1-based adjacent volume element.  Returns a face id >= 0 through *faceId. */
int svfsi_mesh_add_face(void* mesh, int nFe, const int* fIEN, const int* fEl,
                        int* faceId);

/* Zero dVal and dR (LSALLOC, svFSI/LS.f:338-366). */
int svfsi_zero_dev(void* mesh, int dof, double* dVal, double* dR, void* stream);

/* Element loop + DOASSEM into device dVal(dof*dof,nnz), dR(dof,nNo)
 * (accumulates; call svfsi_zero_dev first).  State arrays Ag,Yg,Dg are
 * device (tDof,nNo).  prop_el: optional per-element property (LELAS E_e,
 * FLUID mu_e, NULL = uniform).  quad: 0 = the reference rule (tet4 4-pt,
 * tet10 15-pt), 1 = tet10 4-point rule; for STRUCT, OR one of the
 * SVFSI_STRUCT_* kernel selectors above into quad (default: MFMA64); FLUID
 * on a tet10 mesh takes quad = SVFSI_FLUID_P2P1 (the reference's rules). */
int svfsi_assemble_dev(void* mesh, int phys, const svfsi_eq_params* p,
                       int tDof, const double* dAg, const double* dYg,
                       const double* dDg, const double* dprop_el, int quad,
                       double* dVal, double* dR, void* stream);

/* Face terms (device arrays):
 *   traction: SETBCTRACL (svFSI/SETBC.f:319-416), uniform steady h(3).
 *   backflow: BASSEMNEUBC+BFLUID with h = 0 (EQASSEM.f:90-192,
 *             FLUID.f:1280-1336), outlet backflow stabilization. */
int svfsi_assemble_traction_dev(void* mesh, int faceId, int dof,
                                const double* h, double* dR, void* stream);
int svfsi_assemble_backflow_dev(void* mesh, int faceId,
                                const svfsi_eq_params* p, int tDof,
                                const double* dYg, double* dVal, double* dR,
                                void* stream);

/* ---------------------------------------------------------------------- */
/* Adjoint (no reference counterpart)                                      */
/* ---------------------------------------------------------------------- */

/* Element sensitivity of the LELAS tangent to the per-element modulus:
 * dg[e] = -lam_e^T (dK_e/dE_e) u_e, with dK_e/dE_e the stiffness part at
 * unit modulus (the mass part does not depend on E).  Recomputes element
 * matrices; dlam, du are device (3,nNo), dg device (nEl). */
int svfsi_lelas_dkde_dot_dev(void* mesh, const svfsi_eq_params* p,
                             const double* dlam, const double* du,
                             double* dg, void* stream);

/* Element sensitivity of the FLUID (VMS, tet4) tangent and residual to the
 * per-element viscosity: dg[e] = lam_e^T (dR_e/dmu_e - dK_e/dmu_e du_e),
 * the element term of dL/dmu_e when K du = R and K^T lam = dL/du.  Exact
 * derivative (forward-mode dual numbers through the element routine of
 * svfsi_assemble_dev, including the mu dependence of tau_M, tau_C, tau_B).
 * dAg, dYg: the state of the assembly (tDof,nNo); dmu_el: per-element
 * viscosity or NULL (uniform p->mu); dlam, ddu: (4,nNo); dg: (nEl). */
int svfsi_fluid_dmu_dot_dev(void* mesh, const svfsi_eq_params* p, int tDof,
                            const double* dAg, const double* dYg,
                            const double* dmu_el, const double* dlam,
                            const double* ddu, double* dg, void* stream);

/* Transposed solve K^T lam = rhs with the scaling and Dirichlet mask of
 * fsils_gpu_solve_dev: the values of K^T are formed on the (structurally
 * symmetric) pattern, whose diagonal and hence W are those of K, and the
 * system (W K W)^T y = W rhs is solved like a forward one (GMRES or CG by
 * ls->LS_type; the _mixed_ variant uses the mixed-precision solver of
 * fsils_gpu_solve_mixed_dev).  dVal is the UNSCALED assembled matrix and is
 * not modified; drhs(dof,nNo) in, dlam(dof,nNo) out. */
int fsils_gpu_solve_t_dev(void* lhs, FSILS_lsType* ls, int dof,
                          const double* dVal, const double* drhs,
                          double* dlam, const int* incL, void* stream);
int fsils_gpu_solve_t_mixed_dev(void* lhs, FSILS_lsType* ls, int dof,
                                const double* dVal, const double* drhs,
                                double* dlam, const int* incL, void* stream);

/* Host utility of the polynomial inner preconditioner (no reference
 * counterpart): the harmonic Ritz values of an Arnoldi Hessenberg Hu
 * ((d+1) x d, column-major, leading dimension ld), Leja-ordered and grouped
 * into factors of 1 - a p(a): kind[s] = 2: 1 - coef[2s] a + coef[2s+1] a^2
 * (a conjugate pair or two real roots), kind[s] = 1: 1 - coef[2s] a.
 * Returns the number of factors (<= maxFactors) or a negative code. */
int svfsi_poly_factors(int d, int ld, const double* Hu, double* coef,
                       int* kind, int maxFactors);

/* ---------------------------------------------------------------------- */
/* Profiling hook (HIP events around the SpMV launches of the solvers)     */
/* ---------------------------------------------------------------------- */
int svfsi_prof_enable(int on);
/* kernel: 0 = FP64 SpMV, 1 = FP32 SpMV (including the inner solver's SpMV
 * fused with its multi-dot, and the polynomial preconditioner's steps),
 * 2 = STRUCT element kernels (one per colour), 3 = STRUCT material-point
 * kernel, 4 = the NS solver's Schur-complement product SP = L P - Gt GP
 * (with its partial dot), 5 = FP32 SpMV of the multigrid inner
 * preconditioner's coarse levels (its finest level counts as 1).  Returns
 * launches and total ms. */
int svfsi_prof_get(int kernel, long long* count, double* total_ms);
/* Algorithmic bytes of the profiled launches of `kernel` (values, indices,
 * row pointers, x and y once; for the fused inner step also the basis
 * vectors its multi-dot reads and its partial sums).  For kernel 2 the
 * flops of the B^T Dm B contraction (2 * 30 * 30 * 6 nG per element). */
int svfsi_prof_get_bytes(int kernel, double* bytes);
int svfsi_prof_reset(void);

/* ---------------------------------------------------------------------- */
/* Mesh files (host): the input side of the path                           */
/* ---------------------------------------------------------------------- */
/* Replace READCCNE (svFSI/LOADMSH.f:125-239): the connectivity file has one
 * element per line, "index node_1 .. node_eNoN" (1-based node ids), the
 * coordinates file one node per line, "index x_1 .. x_nsd"; the widths come
 * from the first lines (CheckNoNumbers).  Blank lines are skipped.
 * svfsi_read_ccne fills ien(eNoN, nEl) (1-based, as read) and x(nsd, nNo). */
int svfsi_read_ccne_sizes(const char* connPath, const char* coordPath, int* nEl, int* eNoN, int* nNo,
                          int* nsd);
int svfsi_read_ccne(const char* connPath, const char* coordPath, int nEl, int eNoN, int nNo, int nsd, int* ien,
                    double* x);
/* Replace the face-file branch of READCCNE (LOADMSH.f:196-216): one face
 * element per line, "element index node_1 .. node_feNoN"; fills gE(nFe)
 * and fien(feNoN, nFe) as read. */
int svfsi_read_ebc_sizes(const char* path, int* nFe, int* feNoN);
int svfsi_read_ebc(const char* path, int nFe, int feNoN, int* gE, int* fien);
/* VTK XML meshes and faces, the .vtu/.vtp files svFSI reads
 * (svFSI/VTKXML.f, vtkXMLParser.f90): UnstructuredGrid cells or PolyData
 * polys of one size, 3-component Points, optional PointData
 * "GlobalNodeID" and CellData "GlobalElementID"; DataArrays in ascii,
 * binary (base64) or appended (raw or base64) form, UInt32 or UInt64
 * headers, optionally vtkZLibDataCompressor blocks, little-endian.
 * svfsi_read_vtk fills x(3, nPoints), conn(nodesPerCell, nCells) (0-based,
 * as stored) and, when not NULL, the two id arrays. */
int svfsi_read_vtk_sizes(const char* path, int* nPoints, int* nCells, int* nodesPerCell, int* hasNodeIds,
                         int* hasElementIds);
int svfsi_read_vtk(const char* path, int nPoints, int nCells, int nodesPerCell, double* x, int* conn,
                   int* nodeIds, int* elementIds);
/* Replaces CHECKIEN (svFSI/READMSH.f:1010-1185) for tet4 and tet10 in 3-D,
 * x(3, nNo): bounds and isolated-node checks, then every element whose
 * (x4 - x3) . ((x2 - x1) x (x3 - x2)) is positive (READMSH.f:1131-1141)
 * has nodes 1 and 2 swapped in place (and 6/7, 8/9 for tet10), which makes
 * the isoparametric Jacobian of GNN positive; *nReordered (may be NULL)
 * counts the swapped elements.  A zero sign is an error ("distorted"). */
int svfsi_checkien(int eNoN, int nEl, int nNo, int* ien, const double* x, int* nReordered);

#ifdef __cplusplus
}
#endif
#endif
