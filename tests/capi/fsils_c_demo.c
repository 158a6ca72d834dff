/* A C driver of the svFSILS C API (the prototypes of the reference's
 * svFSILS/FSILS_api_c.h:52-64 and types of FSILS_struct_c.h:60-140), linked
 * against libsvfsi_amd.so in place of svFSILS.
 *
 * Built two ways by the tests:
 *   -DUSE_REFERENCE_HEADER: includes mpi.h and the reference's own FSILS_c.h
 *     first, then svfsi_amd.h, whose fsils_*_ prototypes the compiler then
 *     checks against the reference's (a mismatch is a compile error);
 *   otherwise: svfsi_amd.h alone (where the reference is absent).
 *
 *   fsils_c_demo                 host-only calls: LS defaults, COMMU fields,
 *                                the visible device count
 *   fsils_c_demo fatal           frees an LS that was never created: the
 *                                call stops the program, as the reference's
 *                                STOP "FSILS: FATAL ERROR" does
 *   fsils_c_demo solve IN OUT    reads a system (below), solves it with
 *                                COMMU_CREATE, LHS_CREATE_C, BC_CREATE,
 *                                LS_CREATE and FSILS_SOLVE, writes the result
 *
 * IN (little-endian): int32 nNo, nnz, dof, nDir, LS_type, maxItr, dimKry;
 * float64 relTol, absTol; int32 rowPtr(nNo+1), colPtr(nnz), dirNodes(nDir)
 * (1-based); float64 Val(dof*dof, nnz), R(dof, nNo).
 * OUT: int32 suc, itr; float64 iNorm, fNorm, dB; float64 x(dof, nNo).
 */
#ifdef USE_REFERENCE_HEADER
#include <mpi.h>
#include "FSILS_c.h"
#define MPI_AINT_TYPE MPI_Aint
#else
#define MPI_AINT_TYPE fsils_aint_t
#endif
#include "svfsi_amd.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void rd(FILE* f, void* p, size_t sz, size_t n)
{
    if (fread(p, sz, n, f) != n) {
        fprintf(stderr, "short read\n");
        exit(2);
    }
}

static void host_calls(void)
{
    int types[4] = {LS_TYPE_GMRES, LS_TYPE_CG, LS_TYPE_NS, LS_TYPE_BICGS};
    for (int k = 0; k < 4; ++k) {
        FSILS_lsType ls;
        memset(&ls, 0, sizeof ls);
        fsils_ls_create_(&ls, &types[k], NULL, NULL, NULL, NULL, NULL, NULL, NULL);
        printf("LS %d %d %.3g %d %d %.3g %d %d %.3g %d\n", types[k], ls.foC, ls.RI.relTol, ls.RI.mItr, ls.RI.sD,
               ls.GM.relTol, ls.GM.mItr, ls.GM.sD, ls.CG.relTol, ls.CG.mItr);
        fsils_ls_free_(&ls);
    }
    FSILS_commuType commu;
    memset(&commu, 0, sizeof commu);
    MPI_AINT_TYPE comm = 91;
    fsils_commu_create_(&commu, &comm);
    printf("COMMU %d %d %d %d %d %d %ld\n", commu.foC, commu.masF, commu.master, commu.task, commu.tF,
           commu.nTasks, (long)commu.comm);
    fsils_commu_free_(&commu);
    printf("DEVICES %d\n", svfsi_device_count());
}

static int solve(const char* in, const char* out)
{
    FILE* f = fopen(in, "rb");
    if (!f) return 2;
    int hdr[7];
    double tol[2];
    rd(f, hdr, 4, 7);
    rd(f, tol, 8, 2);
    int nNo = hdr[0], nnz = hdr[1], dof = hdr[2], nDir = hdr[3], LS_type = hdr[4], maxItr = hdr[5],
        dimKry = hdr[6];
    int* rowPtr = malloc(sizeof(int) * (nNo + 1));
    int* colPtr = malloc(sizeof(int) * nnz);
    int* dir = malloc(sizeof(int) * (nDir ? nDir : 1));
    int* gNodes = malloc(sizeof(int) * nNo);
    double* Val = malloc(sizeof(double) * (size_t)dof * dof * nnz);
    double* R = malloc(sizeof(double) * (size_t)dof * nNo);
    rd(f, rowPtr, 4, nNo + 1);
    rd(f, colPtr, 4, nnz);
    rd(f, dir, 4, nDir);
    rd(f, Val, 8, (size_t)dof * dof * nnz);
    rd(f, R, 8, (size_t)dof * nNo);
    fclose(f);
    for (int a = 0; a < nNo; ++a) gNodes[a] = a + 1;

    /* the call sequence of svFSI's INITIALIZE.f:204-208, BAFINI.f:510,
     * READFILES.f:1210 and LS.f:397-398 */
    FSILS_commuType commu;
    memset(&commu, 0, sizeof commu);
    MPI_AINT_TYPE comm = 0;
    fsils_commu_create_(&commu, &comm);
    void* lhs = NULL;
    int nFaces = 1;
    fsils_lhs_create_c_(&lhs, &commu, &nNo, &nNo, &nnz, gNodes, rowPtr, colPtr, &nFaces);
    int faIn = 1, bc = BC_TYPE_Dir;
    fsils_bc_create_(lhs, &faIn, &nDir, &dof, &bc, dir, NULL);
    FSILS_lsType ls;
    memset(&ls, 0, sizeof ls);
    fsils_ls_create_(&ls, &LS_type, &tol[0], &tol[1], &maxItr, dimKry > 0 ? &dimKry : NULL, NULL, NULL, NULL);
    int prec = PRECOND_FSILS;
    fsils_solve_(lhs, &ls, &dof, R, Val, &prec, NULL, NULL);
    fsils_bc_free_(lhs, &faIn);
    fsils_lhs_free_(lhs);
    fsils_ls_free_(&ls);
    fsils_commu_free_(&commu);

    FILE* g = fopen(out, "wb");
    if (!g) return 2;
    int st[2] = {ls.RI.suc, ls.RI.itr};
    double nr[3] = {ls.RI.iNorm, ls.RI.fNorm, ls.RI.dB};
    fwrite(st, 4, 2, g);
    fwrite(nr, 8, 3, g);
    fwrite(R, 8, (size_t)dof * nNo, g);
    fclose(g);
    printf("SOLVE %d %d %.6e %.6e\n", ls.RI.suc, ls.RI.itr, ls.RI.iNorm, ls.RI.fNorm);
    free(rowPtr), free(colPtr), free(dir), free(gNodes), free(Val), free(R);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 1) {
        host_calls();
        return 0;
    }
    if (!strcmp(argv[1], "fatal")) {
        FSILS_lsType ls;
        memset(&ls, 0, sizeof ls);
        fsils_ls_free_(&ls); /* stops the program */
        printf("NOT STOPPED\n");
        return 0;
    }
    if (!strcmp(argv[1], "solve") && argc == 4) return solve(argv[2], argv[3]);
    fprintf(stderr, "usage: fsils_c_demo [fatal | solve IN OUT]\n");
    return 2;
}
