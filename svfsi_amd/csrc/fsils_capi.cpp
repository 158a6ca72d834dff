// fsils_capi.cpp -- the reference's own C entry points of svFSILS
// (Code/Source/svFSILS/FSILS_api_c.h:52-64, types of FSILS_struct_c.h:60-140),
// with their exact by-pointer prototypes, over the fsils_gpu_* entry points
// of this library.  A C or C++ driver written against FSILS_c.h links this
// library instead of svFSILS and runs the solve on the GPU unchanged.
//
// Error behaviour follows the reference: these functions return nothing, and
// misuse prints the reason and stops the program, as STOP "FSILS: FATAL
// ERROR" does (LS.f:60-63, LHS.f:65-69, BC.f:136-150, SOLVE.f:78-81).
// svfsi_fsils_fatal(0) switches the stop off for callers that prefer to poll
// svfsi_fsils_status() and svfsi_last_error() after each call.
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "internal.hpp"

namespace {

int g_fatal = 1;
int g_status = SVFSI_OK;

// Records the status of a reference-API call; a failure stops the program
// unless stops are switched off.
void finish(int rc)
{
    g_status = rc;
    if (rc == SVFSI_OK || !g_fatal) return;
    std::fprintf(stderr, " FSILS: %s\nFSILS: FATAL ERROR\n", svfsi_last_error());
    std::fflush(stderr);
    std::exit(EXIT_FAILURE);
}

// Size and rank of a Fortran MPI communicator handle, through the MPI
// library's Fortran bindings when the process has one loaded (the reference
// calls MPI_COMM_RANK / MPI_COMM_SIZE, COMMU.f:69-70).  Without MPI in the
// process the program is one rank.
void comm_size_rank(int fcomm, int* size, int* rank)
{
    *size = 1;
    *rank = 0;
    using InitFn = void (*)(int*, int*);
    using CommFn = void (*)(int*, int*, int*);
    auto initialized = reinterpret_cast<InitFn>(dlsym(RTLD_DEFAULT, "mpi_initialized_"));
    auto csize = reinterpret_cast<CommFn>(dlsym(RTLD_DEFAULT, "mpi_comm_size_"));
    auto crank = reinterpret_cast<CommFn>(dlsym(RTLD_DEFAULT, "mpi_comm_rank_"));
    if (!initialized || !csize || !crank) return;
    int flag = 0, ierr = 0;
    initialized(&flag, &ierr);
    if (ierr || !flag) return;
    int s = 1, r = 0;
    csize(&fcomm, &s, &ierr);
    if (ierr) return;
    crank(&fcomm, &r, &ierr);
    if (ierr) return;
    *size = s;
    *rank = r;
}

} // namespace

using namespace svfsi;

extern "C" {

int svfsi_fsils_fatal(int on)
{
    const int old = g_fatal;
    g_fatal = on ? 1 : 0;
    return old;
}

int svfsi_fsils_status(void) { return g_status; }

void fsils_commu_create_(FSILS_commuType* commu, fsils_aint_t* commi)
{
    finish(guarded([&] {
        SV_ARG(commu && commi, "commu or commi is NULL");
        if (commu->foC) {
            set_error("COMMU is not free, you may use FSILS_COMMU_FREE to free it (COMMU.f:56-60)");
            return SVFSI_ERR_STATE;
        }
        // The Fortran routine reads commi as INTEGER(KIND=LSIP) (COMMU.f:53):
        // the low 4 bytes of the MPI_Aint on a little-endian host.
        int handle = 0;
        std::memcpy(&handle, commi, sizeof(int));
        int size = 1, rank = 0;
        comm_size_rank(handle, &size, &rank);
        commu->foC = 1;
        commu->comm = (fsils_aint_t)handle;
        commu->nTasks = size;
        commu->task = rank;
        commu->master = 0;
        commu->masF = rank == 0;
        commu->tF = rank + 1;
        return SVFSI_OK;
    }));
}

void fsils_commu_free_(FSILS_commuType* commu)
{
    finish(guarded([&] {
        SV_ARG(commu, "commu is NULL");
        if (!commu->foC) {
            set_error("COMMU is not created yet to be freed (COMMU.f:92)");
            return SVFSI_ERR_STATE;
        }
        commu->foC = 0;
        return SVFSI_OK;
    }));
}

void fsils_lhs_create_c_(void** lhs, FSILS_commuType* commu, int* gnNo, int* nNo, int* nnz, int* gNodes,
                         int* rowPtr, int* colPtr, int* nFaces)
{
    finish(guarded([&] {
        SV_ARG(lhs && commu && gnNo && nNo && nnz && nFaces, "NULL argument");
        if (!commu->foC) {
            set_error("COMMU is not created (FSILS_COMMU_CREATE comes first)");
            return SVFSI_ERR_STATE;
        }
        if (commu->nTasks != 1) {
            set_error("communicator of %d ranks: this library solves on one GPU, one rank (the multi-rank "
                      "reordering of LHS.f:113-288 is not built)",
                      commu->nTasks);
            return SVFSI_ERR_NOTSUP;
        }
        return fsils_gpu_lhs_create(lhs, *gnNo, *nNo, *nnz, gNodes, rowPtr, colPtr, *nFaces);
    }));
}

void fsils_lhs_free_(void* lhs)
{
    finish(guarded([&] {
        SV_ARG(lhs, "LHS is not created yet to be freed (LHS.f:301-304)");
        return fsils_gpu_lhs_free(lhs);
    }));
}

void fsils_ls_create_(FSILS_lsType* ls, int* LS_type, double* relTol, double* absTol, int* maxItr, int* dimKry,
                      double* relTolIn, double* absTolIn, int* maxItrIn)
{
    finish(guarded([&] {
        SV_ARG(ls && LS_type, "ls or LS_type is NULL");
        return fsils_gpu_ls_create(ls, *LS_type, relTol, absTol, maxItr, dimKry, relTolIn, absTolIn, maxItrIn);
    }));
}

void fsils_ls_free_(FSILS_lsType* ls)
{
    finish(guarded([&] {
        SV_ARG(ls, "ls is NULL");
        if (!ls->foC) {
            set_error("LS is not created yet to be freed (LS.f:125-128)");
            return SVFSI_ERR_STATE;
        }
        ls->foC = 0;
        return SVFSI_OK;
    }));
}

void fsils_bc_create_(void* lhs, int* faIn, int* nNo, int* dof, int* BC_type, int* gNodes, double* Val)
{
    finish(guarded([&] {
        SV_ARG(lhs && faIn && nNo && dof && BC_type, "NULL argument");
        SV_ARG(gNodes || *nNo == 0, "gNodes is NULL");
        return fsils_gpu_bc_create(lhs, *faIn, *nNo, *dof, *BC_type, gNodes, Val);
    }));
}

void fsils_bc_free_(void* lhs, int* faIn)
{
    finish(guarded([&] {
        SV_ARG(lhs && faIn, "NULL argument");
        return fsils_gpu_bc_free(lhs, *faIn);
    }));
}

void fsils_solve_(void* lhs, FSILS_lsType* ls, int* dof, double* Ri, double* Val, int* prec, int* incL,
                  double* res)
{
    finish(guarded([&] {
        SV_ARG(lhs && ls && dof && Ri && Val && prec, "NULL argument");
        return fsils_gpu_solve(lhs, ls, *dof, Ri, Val, *prec, incL, res);
    }));
}

} // extern "C"
