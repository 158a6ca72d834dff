// internal.hpp -- shared definitions of libsvfsi_amd (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "../../include/svfsi_amd.h"

namespace svfsi {

void set_error(const char* fmt, ...);

#define SV_HIP(x)                                                              \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) {                                                \
            ::svfsi::set_error("%s:%d %s: %s", __FILE__, __LINE__, #x,         \
                               hipGetErrorString(e_));                         \
            return SVFSI_ERR_HIP;                                              \
        }                                                                      \
    } while (0)

#define SV_TRY(x)                                                              \
    do {                                                                       \
        int r_ = (x);                                                          \
        if (r_ != SVFSI_OK) return r_;                                         \
    } while (0)

#define SV_ARG(cond, ...)                                                      \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::svfsi::set_error(__VA_ARGS__);                                   \
            return SVFSI_ERR_ARG;                                              \
        }                                                                      \
    } while (0)

// Runs the body of an extern "C" entry point and turns any C++ exception
// (std::bad_alloc, std::length_error of a container) into an error code, so
// that none crosses the C ABI.
template <class F>
int guarded(F&& body) noexcept
{
    try {
        return body();
    } catch (const std::exception& e) {
        set_error("internal error: %s", e.what());
    } catch (...) {
        set_error("internal error: unknown exception");
    }
    return SVFSI_ERR_ARG;
}

int require_device();

// Device buffer owned by the library (hipMalloc'd, freed in destructor).
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int ensure(size_t nbytes) {
        if (nbytes <= bytes && p) return SVFSI_OK;
        release();
        if (nbytes == 0) nbytes = 8;
        SV_HIP(hipMalloc(&p, nbytes));
        bytes = nbytes;
        return SVFSI_OK;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

struct Face {
    bool foC = false;
    int nNo = 0, dof = 0, bGrp = BC_TYPE_Dir;
    std::vector<int> glob;   // 0-based nodes
    std::vector<double> val; // (dof, nNo)
};

// Pinned, double-buffered host mirror of a small device state: the host
// polls the state of the previous chunk of iterations while the next one
// runs.  Allocated once per workspace.
struct PinnedPoll {
    void* h[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    size_t bytes = 0;
    PinnedPoll() = default;
    PinnedPoll(const PinnedPoll&) = delete;
    PinnedPoll& operator=(const PinnedPoll&) = delete;
    ~PinnedPoll();
    int ensure(size_t b);
    int post(int q, const void* dsrc, size_t n, hipStream_t s);
    int wait(int q);
};

// Instantiated HIP graphs of fixed kernel sequences (one per key), captured
// on a private stream and replayed on the caller's stream.  `sig` fingerprints
// every pointer and size baked into the captured launches; a different
// signature drops all entries.
struct GraphCache {
    struct Entry {
        int key;
        hipGraphExec_t exec;
    };
    std::vector<Entry> e;
    std::vector<uintptr_t> sig;
    hipStream_t cap = nullptr;
    GraphCache() = default;
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    ~GraphCache();
    void clear();
};

struct AmgHier; // aggregation multigrid hierarchy (amg.hip)

// Krylov workspace, grown on demand and reused across solves.
struct Work {
    GraphCache graphs;   // mixed solve: one graph per outer step
    GraphCache gmGraphs; // FP64 GMRES: one graph per chunk of Arnoldi steps
    GraphCache cgGraphs; // CG: one graph for a chunk of iterations
    GraphCache csGraphs; // NS solver's Schur-complement CG: likewise
    PinnedPoll poll;
    DBuf W, X, R0, tmp, basis, basis32, Z32, val32, x32, y32;
    DBuf partial, state, hess;
    DBuf zdir, opartial, ostate, ohess; // flexible GMRES: FP32 directions, outer state
    DBuf mask;
    DBuf rcs, W1;  // PRECOND_RCS: row/column factors, partial maxima; W1 (row scaling)
    DBuf ns;       // NS solver: the split blocks, its vectors and directions
    DBuf bicgs;    // BiCGStab: P, Rh, X, V, S, T and the dot partials
    DBuf valT; // K^T values for transposed solves
    DBuf dinv32; // inverse diagonal blocks (inner precision) of the inner block-Jacobi
    DBuf valIn;  // FP64 nested solve: the block-Jacobi scaled inner matrix
    std::shared_ptr<AmgHier> amg; // inner multigrid of the nested solve (pattern part kept)
    DBuf polyCoef, polyVec; // polynomial preconditioner: factors, work vectors
    int maskDof = -1;
    std::vector<int> maskInc;
    int maskVersion = -1;
};

// Coupled Neumann faces of the solve in progress (BC_TYPE_Neu faces with
// res(faIn) != 0, svFSILS/SOLVE.f:78-86), concatenated: face f holds entries
// off[f] .. off[f+1]-1 of glob (0-based node), val (the face vector, stride
// dof, fdof[f] components) and vm (val scaled by the preconditioner,
// PRECOND.f:133-140).  coef[f] = res, coef[nC+f] = -res/(1 + res nS) (BCPRE).
struct CplSet {
    int nC = 0, T = 0, dof = 0;
    bool disjoint = true;  // no node on two faces: one launch, in place
    bool active = false;   // set for the duration of a solve
    unsigned gen = 0;      // bumped when the device layout changes
    std::vector<int> ids;  // lhs face of each coupled face (layout key)
    int keyDof = 0, keyVersion = -1;
    DBuf off, glob, fdof, val, vm, coef, tmp;
    std::vector<double> hres;
};

struct Lhs {
    int gnNo = 0, nNo = 0, nnz = 0, nFaces = 0;
    DBuf rowPtr, colPtr, diagPtr, blockRow, trans;
    std::vector<int> hRowPtr, hColPtr; // 0-based host copies
    std::vector<Face> faces;
    int facesVersion = 0;
    bool symmetric = false; // every block (i,j) has a (j,i) partner: trans[] valid
    // Row tiles of the CSR-stream block products: tile b is rows
    // tiles[t][b] .. tiles[t][b+1]-1, at most 256 (t = 0), 512 (t = 1) or
    // 1024 (t = 2) blocks together; nTiles[t] = 0 when some row holds more
    DBuf tiles[3];
    int nTiles[3] = {0, 0, 0};
    CplSet cpl;
    Work ws;
};

struct Mesh {
    Lhs* lhs = nullptr;
    int nNo = 0, nEl = 0, eNoN = 0;
    DBuf x;          // (3, nNo)
    DBuf ien;        // (eNoN, nEl) colour-sorted, 0-based
    DBuf elemId;     // original element id of each colour-sorted slot
    DBuf emap;       // (eNoN*eNoN, nEl) colour-sorted nnz index of block (a,b)
    std::vector<int> colorOff; // nColors+1 offsets into the colour-sorted list
    struct FaceSet {
        int nFe = 0;
        DBuf fien;   // (3, nFe) 0-based
        DBuf fel;    // (nFe) 0-based volume element
        DBuf inner;  // (nFe) first element node not on the face
        std::vector<int> colorOff;
        DBuf fmap;   // (9, nFe) nnz index of face block (a,b)
        DBuf order;  // colour-sorted face ids
    };
    std::vector<FaceSet*> faceSets;
    DBuf rule10[2];  // TET10 quadrature tables (15-point, 4-point), built on first use
    DBuf ruleTH;     // Taylor-Hood FLUID quadrature (19 points), built on first use
    DBuf midNodes;   // tet10: the nodes that are a mid-edge node of some element
    int nMid = 0;
    int rule10nG[2] = {0, 0};
    DBuf srec;       // TET10 material-point records (nEl x nG x 97 doubles), grown on demand
    // Gather assembly: element block (slot p, a, b) is stored at position
    // gSlot[p*eNoN^2 + a*eNoN + b] of ke, so that the contributions to nnz
    // block j are ke positions gPtr[j] .. gPtr[j+1]-1, in element-id order
    // (the order of DOASSEM's element loop); likewise element row (p, a) at
    // rSlot[p*eNoN + a] of re, node i's rows at rPtr[i] .. rPtr[i+1]-1.
    DBuf gPtr, gSlot, rPtr, rSlot;
    DBuf ke, re;     // element blocks / element residual rows, grown on demand
    ~Mesh() {
        for (auto* f : faceSets) delete f;
    }
};

// Profiling of SpMV launches (HIP events).
struct ProfPending {
    hipEvent_t a, b;
    int k;
    double bytes;
};
struct Prof {
    bool on = false;
    // slots: 0 FP64 SpMV, 1 FP32 SpMV (inner solver), 2 STRUCT element
    // kernels, 3 STRUCT material-point kernel, 4 NS Schur-complement product,
    // 5 FP32 SpMV of the multigrid's coarse levels
    static constexpr int kSlots = 6;
    long long count[kSlots] = {};
    double ms[kSlots] = {};
    double bytes[kSlots] = {}; // algorithmic bytes (slot 2: MFMA flops) of the timed launches
    std::vector<ProfPending> pending;
};
Prof& prof();
void prof_resolve();

// Kernel-accurate timing of a profiled launch: the events are attached to
// the dispatch itself (hipExtLaunchKernelGGL), so they bracket the kernel's
// execution and not the launch gaps around it.  Resolved lazily by
// prof_resolve() so that profiling adds no host synchronisation.
struct ProfScope {
    hipEvent_t a = nullptr, b = nullptr;
    int k;
    double bytes;
    explicit ProfScope(int kernel, double algBytes = 0.0) : k(kernel), bytes(algBytes) {
        if (prof().on) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
        }
    }
    ~ProfScope() {
        if (a) prof().pending.push_back({a, b, k, bytes});
    }
};

#define SV_LAUNCH(ps, kern, grid, block, shm, st, ...)                                           \
    do {                                                                                          \
        if ((ps).a)                                                                               \
            hipExtLaunchKernelGGL(kern, grid, block, shm, st, (ps).a, (ps).b, 0, __VA_ARGS__);    \
        else                                                                                      \
            hipLaunchKernelGGL(kern, grid, block, shm, st, __VA_ARGS__);                          \
    } while (0)


// ---- launchers (device pointers; stream may be null) ----
int launch_spmv(const Lhs& L, int dof, const double* K, const double* U,
                double* KU, hipStream_t s, const int* stop = nullptr);
int launch_spmv_f32(const Lhs& L, int dof, const float* K, const float* U,
                    float* KU, hipStream_t s, const int* stop = nullptr,
                    const float* dinv = nullptr, int profSlot = 1);
// The inner (preconditioner) kernels of the nested solve, instantiated for
// IT = float (the mixed solve) and IT = double (the same algorithm in FP64).
template <typename OT>
int launch_blockdiag_inv(const Lhs& L, int dof, const double* val, OT* dinv, hipStream_t s);
// number of unmasked dofs of the scaled matrix with a zero diagonal entry
// (dCount: one device int of scratch)
int count_zero_diag(const Lhs& L, int dof, const double* val, const double* W, int* dCount, int& hCount,
                    hipStream_t s);
// 1 when the SpMV-shaped kernels map blocks XCD-contiguously (xcd_block)
int xcd_map();
// the same from FP32 values (the coarse levels of the aggregation hierarchy)
int launch_blockdiag_inv_f32(const Lhs& L, int dof, const float* val, float* dinv, hipStream_t s);
template <typename OT>
int launch_bjscale(const Lhs& L, int dof, const double* K, const OT* dinv, OT* B, hipStream_t s);
template <typename IT>
int launch_poly(const Lhs& L, int dof, int mode, const IT* K, const IT* dinv, const IT* x, const IT* prodOld, IT* w,
                IT* y, IT* out, const double* coef, int step, hipStream_t s, const int* stop);
template <typename IT>
int launch_poly_last(size_t n, const IT* prod, IT* y, const double* coef, int step, hipStream_t s, const int* stop);
template <typename IT>
int launch_rand_masked(size_t n, unsigned seed, const double* W, IT* v, hipStream_t s);
int poly_factors(int d, int ld, const double* Hu, std::vector<double>& coef, std::vector<int>& kind);
template <typename IT>
int launch_blockdiag_apply(const Lhs& L, int dof, const IT* dinv, IT* x, hipStream_t s, const int* stop);
int launch_spmv_t(const Lhs& L, int dof, const double* K, const double* U,
                  double* KU, hipStream_t s, const int* stop = nullptr);
int launch_transpose_blocks(const Lhs& L, int dof, const double* V, double* VT, hipStream_t s);
// r = b - K x (or b - K^T x) with double-double row sums, rounded once
int launch_resid_dd(const Lhs& L, int dof, const double* K, const double* x, const double* b, double* r,
                    bool transpose, hipStream_t s);
int launch_dot(int n, const double* U, const double* V, double* partial,
               double* result, hipStream_t s);
int build_mask(Lhs& L, int dof, const int* incL, hipStream_t s);
// PRECONDRCS: Val, R scaled in place; W1 row, W2 column factors (the
// solution of the scaled system is unscaled with W2); optional FP32 copy.
int precond_rcs(Lhs& L, int dof, double* Val, double* R, double* W1, double* W2, float* Val32, const int* incL,
                hipStream_t s);
int precond_diag(Lhs& L, int dof, double* Val, double* R, double* W,
                 float* Val32, const int* incL, hipStream_t s);
// PRECONDDIAG with a Dirichlet mask the caller provides (device, (dof,nNo)).
int precond_diag_masked(const Lhs& L, int dof, double* Val, double* R, double* W, float* Val32,
                        const double* mask, hipStream_t s);
// The Dirichlet mask of build_mask, built on the device into `mask` with no
// allocation; `staging` holds one face's nodes and values at a time.
int build_mask_into(const Lhs& L, int dof, const int* incL, double* mask, void* staging, size_t stagingBytes,
                    hipStream_t s);
int launch_fill(size_t n, double v, double* x, hipStream_t s);
// lanes per block row of the dof-3 SpMV kernels for this pattern
int spmv3_lanes(const Lhs& L);
// the CSR-stream row tiles of L (from hRowPtr; rows must be non-empty)
int build_row_tiles(Lhs& L);
// the row tiles the dof-3/4 products use for this pattern and precision, or
// nullptr (then the row-per-lane-group kernels run); T: blocks per tile
const int* stream_tiles(const Lhs& L, int dof, bool fp64, int& T, int& nT);
int launch_scale_vec(size_t n, const double* W, double* R, hipStream_t s);

int cg_solve(Lhs& L, FSILS_subLsType& ls, int dof, const double* K,
             double* R, hipStream_t s);
// CGRADV's device buffers: X, P, KP (n each), the partial sums
// (kCgPartials doubles) and the iteration state (kCgStateBytes).
constexpr size_t kCgPartials = 1024 + 512 + 16;
constexpr size_t kCgStateBytes = 128;
struct CgBufs {
    double *X, *P, *KP, *part;
    void* state;
};
// CGRADV on caller buffers.  graphs: replay chunks of iterations as one
// captured graph (cached in L.ws); pinned: poll the state through L.ws's
// pinned mirror one chunk behind.  With neither, nothing is allocated: the
// host reads the state with a synchronous copy after each chunk.
int cg_core(Lhs& L, FSILS_subLsType& ls, int dof, const double* K, double* R, hipStream_t s, const CgBufs& b,
            bool graphs, bool pinned);
// BICGSV (svFSILS/BICGS.f:115-180): R in (scaled RHS), out (solution).
int bicgs_solve(Lhs& L, FSILS_subLsType& ls, int dof, const double* K, double* R, hipStream_t s);
// CGRAD_SCHUR (svFSILS/CGRAD.f:51-123) on the device: (L - Gt G) x = R with
// G (3, nnz), Gt (3, nnz), Lm (nnz); R in: right-hand side, out: solution.
// X, P, P2, SP (nNo) and GP (3, nNo) are work vectors; ls.itr and ls.callD
// accumulate, as the reference's do.
int cg_schur_solve(Lhs& L, FSILS_subLsType& ls, const double* Gt, const double* G, const double* Lm, double* R,
                   double* X, double* P, double* P2, double* SP, double* GP, hipStream_t s);
// cpl: coupled faces in the operator, 0 none, 1 ADDBCMUL ADD (GMRESV),
// 2 ADD then PRE (the NS solver's GMRES; the caller applies PRE to R)
int gmres_solve(Lhs& L, FSILS_subLsType& ls, int dof, const double* K,
                double* R, hipStream_t s, bool transpose, int cpl = 0);
int mixed_solve(Lhs& L, FSILS_subLsType& ls, int dof, const double* K,
                float* K32, double* R, hipStream_t s);
// Aggregation multigrid (amg.hip): greedy aggregation of a CSR graph
// (0-based) into aggregates of at most cap nodes; set-up of the hierarchy
// and its values for one solve; one V-cycle z = M^-1 q (FP32).
int amg_aggregate(int n, const int* rp, const int* cp, int cap, std::vector<int>& agg);
int amg_setup(Lhs& L, int dof, const double* K64, const float* K32, hipStream_t s);
int amg_apply(Lhs& L, int dof, const float* q, float* z, hipStream_t s);
int amg_levels(const Lhs& L, std::vector<int>& nodes);
// The same nested solver with the inner preconditioner in FP64.
int nested_solve_fp64(Lhs& L, FSILS_subLsType& ls, int dof, const double* K, double* R, hipStream_t s);

double wall_seconds();
// Coupled Neumann faces: cpl_setup selects the faces of this solve (none:
// L.cpl.active = false) and uploads their layout; cpl_scale forms vm = val W
// and the BCPRE coefficients after the diagonal preconditioner;
// cpl_apply is ADDBCMUL (svFSILS/ADDBCMUL.f:53-114): Y += coef v (v . X) per
// face, op 0 ADD, 1 PRE, on vectors of vd components per node.  X may equal
// Y (PRE in place).  stop: optional device flag that skips the launch.
int cpl_setup(Lhs& L, int dof, const int* incL, const double* res, hipStream_t s);
int cpl_scale(Lhs& L, const double* W, hipStream_t s);
int cpl_apply(Lhs& L, int op, int vd, const double* X, double* Y, hipStream_t s, const int* stop = nullptr);
// NSSOLVER (svFSILS/NSSOLVER.f:52-383) on the scaled system, dof 4: Ri in
// (right-hand side), out (solution); ls.RI, ls.GM, ls.CG, Resm, Resc set.
int ns_solve(Lhs& L, FSILS_lsType& ls, int dof, const double* Val, double* Ri, hipStream_t s);

} // namespace svfsi
