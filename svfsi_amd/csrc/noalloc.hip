// noalloc.hip -- FSILS_SOLVE with CGRADV and diagonal scaling that allocates
// nothing: a variant of fsils_gpu_solve_dev (LS_TYPE_CG, PRECOND_FSILS) for
// callers that cannot allocate device memory.  Every device buffer the solve
// needs beyond its inputs and outputs (the scaling W, the Dirichlet mask, the
// CG vectors X, P, KP, the partial sums, the iteration state, and a staging
// area for the face lists) lives in one file-scope device array of fixed size.
// Each GPU holds its own copy of that array for the life of the process; its
// address on a device is looked up once, with that device current, and only
// that address is handed to kernels and copies.
#include <cstring>
#include <mutex>

#include "internal.hpp"

constexpr size_t kScratchBytes = size_t(64) << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call for that device.
int scratch_ptr(void** out)
{
    int dev = -1;
    SV_HIP(hipGetDevice(&dev));
    SV_ARG(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d devices supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        SV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SV_ARG(p != nullptr, "no device address for the scratch array");
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return SVFSI_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

} // namespace

using namespace svfsi;

extern "C" {

size_t svfsi_noalloc_scratch_bytes(void) { return kScratchBytes; }

int fsils_gpu_solve_cg_noalloc_dev(void* lhs, FSILS_lsType* ls, int dof, double* dRi, double* dVal,
                                   const int* incL, void* stream)
{
    return guarded([&]() -> int {
        SV_ARG(lhs && ls && dRi && dVal, "NULL argument");
        SV_ARG(dof >= 2 && dof <= 4, "dof %d not supported (2..4)", dof);
        if (ls->LS_type != LS_TYPE_CG) {
            set_error("the allocation-free solve is CGRADV only (LS_type %d given, LS_TYPE_CG = %d)", ls->LS_type,
                      LS_TYPE_CG);
            return SVFSI_ERR_NOTSUP;
        }
        Lhs& L = *static_cast<Lhs*>(lhs);
        const size_t n = (size_t)L.nNo * dof;
        // W, mask, X, P, KP: n doubles each; partials; state.  The face
        // staging area reuses X..KP, which are free until the CG starts.
        const size_t vec = align256(n * sizeof(double));
        const size_t need = 5 * vec + align256(kCgPartials * sizeof(double)) + align256(kCgStateBytes);
        if (need > kScratchBytes) {
            set_error("system of %zu unknowns needs %zu bytes of scratch; the allocation-free solve holds %zu "
                      "(about %zu unknowns)",
                      n, need, kScratchBytes, (kScratchBytes - 4096) / (5 * sizeof(double)));
            return SVFSI_ERR_ARG;
        }
        SV_TRY(require_device());
        void* base = nullptr;
        SV_TRY(scratch_ptr(&base));
        char* p = static_cast<char*>(base);
        double* W = reinterpret_cast<double*>(p);
        double* mask = reinterpret_cast<double*>(p + vec);
        CgBufs b{};
        b.X = reinterpret_cast<double*>(p + 2 * vec);
        b.P = reinterpret_cast<double*>(p + 3 * vec);
        b.KP = reinterpret_cast<double*>(p + 4 * vec);
        b.part = reinterpret_cast<double*>(p + 5 * vec);
        b.state = p + 5 * vec + align256(kCgPartials * sizeof(double));
        hipStream_t s = static_cast<hipStream_t>(stream);

        SV_TRY(build_mask_into(L, dof, incL, mask, b.X, 3 * vec, s));
        SV_TRY(precond_diag_masked(L, dof, dVal, dRi, W, nullptr, mask, s));
        SV_TRY(cg_core(L, ls->RI, dof, dVal, dRi, s, b, false, false));
        SV_TRY(launch_scale_vec(n, W, dRi, s));
        SV_HIP(hipStreamSynchronize(s));
        return SVFSI_OK;
    });
}

} // extern "C"
