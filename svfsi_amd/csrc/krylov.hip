// krylov.hip -- device-resident CGRADV / GMRESV for gfx950.
//
// The iteration scalars (alpha, err, Hessenberg column, Givens rotations,
// convergence flag) live in device memory and are produced by the last
// workgroup of each reduction kernel, so an iteration issues no host
// round trip: CG is three kernels per iteration, GMRES three kernels per
// Arnoldi step (SpMV; multi-dot + Pythagorean norm + Givens; update).  The
// host enqueues a chunk of iterations and polls a pinned copy of the state
// of the previous chunk; kernels after convergence exit at their first
// instruction.  All reductions are fixed-order, so results are bitwise
// reproducible run to run.
#include "internal.hpp"
#include "devutil.hpp"
#include "stream.hpp"

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace svfsi {

double wall_seconds()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ===========================================================================
// CGRADV (svFSILS/CGRAD.f:184-242)
// ===========================================================================
struct CgState {
    double err, errO, alpha, eps;
    int done, suc, itr, mItr;
    unsigned int counter;
    int pad;
    double e[2]; // |R|^2 of the current and the next iteration, by iteration parity
};

// Every reduction ends at a kernel boundary (no last-block fences, which on
// a multi-XCD part write back L2): a kernel writes per-block partials and
// each block of the next kernel sums them in the same fixed order, so all
// blocks hold bitwise the same scalar.

__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int m, double* lds)
{
    double t = 0.0;
    for (int q = threadIdx.x; q < m; q += 256) t += part[q];
    return block_sum<256>(t, lds);
}

// K P for the block rows of this block (grid-stride) and per-block
// partials of <P, K P>.
template <int DOF>
__global__ __launch_bounds__(256) void k_cg_spmv_dot(int nNo, const int* __restrict__ rowPtr,
                                                     const int* __restrict__ colPtr,
                                                     const double* __restrict__ val,
                                                     const double* __restrict__ p,
                                                     double* __restrict__ kp,
                                                     double* __restrict__ partial,
                                                     const CgState* __restrict__ st)
{
    if (st->done) return;
    constexpr int DD = DOF * DOF;
    constexpr int G = kWave / DD;
    __shared__ double lds[4];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int g = lane / DD;
    const int r = lane - g * DD;
    const int i = r / DOF;
    const int j = r - i * DOF;
    double contrib = 0.0;
    for (int rb = blockIdx.x; rb * 4 * G < nNo; rb += gridDim.x) {
        const int row = (rb * 4 + wave) * G + g;
        const bool active = (g < G) && (row < nNo);
        double acc = 0.0;
        if (active) {
            int k = rowPtr[row];
            const int e = rowPtr[row + 1];
            double a0 = 0.0, a1 = 0.0;
            for (; k + 1 < e; k += 2) {
                const int c0 = colPtr[k], c1 = colPtr[k + 1];
                a0 += val[(size_t)k * DD + r] * p[(size_t)c0 * DOF + j];
                a1 += val[(size_t)(k + 1) * DD + r] * p[(size_t)c1 * DOF + j];
            }
            if (k < e) a0 += val[(size_t)k * DD + r] * p[(size_t)colPtr[k] * DOF + j];
            acc = a0 + a1;
        }
        if constexpr (DOF == 4) {
            acc += __shfl_xor(acc, 1, kWave);
            acc += __shfl_xor(acc, 2, kWave);
        } else if constexpr (DOF == 2) {
            acc += __shfl_xor(acc, 1, kWave);
        } else {
            double t = 0.0;
#pragma unroll
            for (int q = 1; q < DOF; ++q) t += __shfl_down(acc, q, kWave);
            acc += t;
        }
        if (active && j == 0) {
            kp[(size_t)row * DOF + i] = acc;
            contrib += p[(size_t)row * DOF + i] * acc;
        }
    }
    const double sb = block_sum<256>(contrib, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = sb;
}

// dof 2 / 3: the value-streaming layout of the SpMV (LPR lanes per block
// row reading the row's contiguous values in coalesced strides), grid-stride
// over rows, per-block partials of <P, K P>.
template <int DOF, int LPR>
__global__ __launch_bounds__(256) void k_cg_spmv_dot_vs(int nNo, const int* __restrict__ rowPtr,
                                                        const int* __restrict__ colPtr,
                                                        const double* __restrict__ val,
                                                        const double* __restrict__ p,
                                                        double* __restrict__ kp,
                                                        double* __restrict__ partial,
                                                        const CgState* __restrict__ st)
{
    if (st->done) return;
    constexpr int DD = DOF * DOF;
    constexpr int RPB = 256 / LPR;
    __shared__ double lds[4];
    const int l = threadIdx.x % LPR;
    double contrib = 0.0;
    for (int rb = blockIdx.x; rb * RPB < nNo; rb += gridDim.x) {
        const int row = rb * RPB + threadIdx.x / LPR;
        double acc[DOF];
#pragma unroll
        for (int i = 0; i < DOF; ++i) acc[i] = 0.0;
        if (row < nNo) {
            const int k0 = rowPtr[row];
            const int nv = (rowPtr[row + 1] - k0) * DD;
            const double* vr = val + (size_t)k0 * DD;
            const int* cr = colPtr + k0;
            for (int v = l; v < nv; v += LPR) {
                const int k = v / DD, e = v - k * DD;
                const double a = vr[v] * p[(size_t)cr[k] * DOF + (e % DOF)];
#pragma unroll
                for (int i = 0; i < DOF; ++i) acc[i] += (e / DOF == i) ? a : 0.0;
            }
        }
#pragma unroll
        for (int i = 0; i < DOF; ++i)
#pragma unroll
            for (int o = LPR >> 1; o > 0; o >>= 1) acc[i] += __shfl_xor(acc[i], o, kWave);
        if (row < nNo && l < DOF) {
            double out = acc[0];
#pragma unroll
            for (int i = 1; i < DOF; ++i)
                if (l == i) out = acc[i];
            kp[(size_t)row * DOF + l] = out;
            contrib += p[(size_t)row * DOF + l] * out;
        }
    }
    const double sb = block_sum<256>(contrib, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = sb;
}

// dof 3 on the SpMV's dof-4 lane mapping (k_bspmv3r, H = 4): 16 lanes per
// block row, lane (h, i) takes blocks s + h, s + h + 4, ... and block row
// i < 3; grid-stride over rows, per-block partials of <P, K P>.
__global__ __launch_bounds__(256) void k_cg_spmv_dot3r(int nNo, const int* __restrict__ rowPtr,
                                                       const int* __restrict__ colPtr,
                                                       const double* __restrict__ val,
                                                       const double* __restrict__ p,
                                                       double* __restrict__ kp,
                                                       double* __restrict__ partial,
                                                       const CgState* __restrict__ st)
{
    if (st->done) return;
    constexpr int H = 4, RPB = 256 / (4 * H);
    __shared__ double lds[4];
    const int l = threadIdx.x % (4 * H);
    const int h = l >> 2, i = l & 3;
    double contrib = 0.0;
    for (int rb = blockIdx.x; rb * RPB < nNo; rb += gridDim.x) {
        const int row = rb * RPB + threadIdx.x / (4 * H);
        double a0 = 0.0, a1 = 0.0;
        if (row < nNo && i < 3) {
            const int e = rowPtr[row + 1];
            int k = rowPtr[row] + h;
            for (; k + H < e; k += 2 * H) {
                const double* ua = val + (size_t)k * 9 + i * 3;
                const double* ub = val + (size_t)(k + H) * 9 + i * 3;
                const double* xa = p + (size_t)colPtr[k] * 3;
                const double* xb = p + (size_t)colPtr[k + H] * 3;
                a0 += ua[0] * xa[0] + ua[1] * xa[1] + ua[2] * xa[2];
                a1 += ub[0] * xb[0] + ub[1] * xb[1] + ub[2] * xb[2];
            }
            if (k < e) {
                const double* ua = val + (size_t)k * 9 + i * 3;
                const double* xa = p + (size_t)colPtr[k] * 3;
                a0 += ua[0] * xa[0] + ua[1] * xa[1] + ua[2] * xa[2];
            }
        }
        double acc = a0 + a1;
#pragma unroll
        for (int o = 4; o < 4 * H; o <<= 1) acc += __shfl_xor(acc, o, kWave);
        if (row < nNo && h == 0 && i < 3) {
            kp[(size_t)row * 3 + i] = acc;
            contrib += p[(size_t)row * 3 + i] * acc;
        }
    }
    const double sb = block_sum<256>(contrib, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = sb;
}

// The same on the row tiles (stream_tile): workgroups stride over the tiles,
// one partial of <P, K P> per workgroup.
template <int T>
__global__ __launch_bounds__(256) void k_cg_sspmv_dot3(int nTiles, const int* __restrict__ tile,
                                                       const int* __restrict__ rowPtr,
                                                       const int* __restrict__ colPtr,
                                                       const double* __restrict__ val,
                                                       const double* __restrict__ p, double* __restrict__ kp,
                                                       double* __restrict__ partial,
                                                       const CgState* __restrict__ st)
{
    if (st->done) return;
    __shared__ double lds[4];
    double contrib = 0.0;
    for (int b = blockIdx.x; b < nTiles; b += gridDim.x)
        stream_tile<double, 3, T, false>(b, tile, rowPtr, colPtr, val, p, true,
                                         [&](int row, int i, double v, double, double, double, double) {
                                             const size_t q = (size_t)row * 3 + i;
                                             kp[q] = v;
                                             contrib += p[q] * v;
                                         });
    const double sb = block_sum<256>(contrib, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = sb;
}

// alpha = |R|^2 / <P, K P> ; X += alpha P ; R -= alpha KP ; partials of |R|^2
// (CGRAD.f:216-226).  par: parity of the iteration.
__global__ __launch_bounds__(256) void k_cg_update(size_t n, double* __restrict__ x,
                                                   double* __restrict__ rr,
                                                   const double* __restrict__ p,
                                                   const double* __restrict__ kp,
                                                   const double* __restrict__ part1, int m1,
                                                   double* __restrict__ part2,
                                                   CgState* __restrict__ st, int par)
{
    if (st->done) return;
    __shared__ double lds[4];
    const double pkp = sum_partials(part1, m1, lds);
    const double alpha = st->e[par] / pkp;
    double s = 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        x[q] = x[q] + alpha * p[q];
        const double rv = rr[q] + (-alpha) * kp[q];
        rr[q] = rv;
        s += rv * rv;
    }
    s = block_sum<256>(s, lds);
    if (threadIdx.x == 0) {
        part2[blockIdx.x] = s;
        if (blockIdx.x == 0) st->alpha = alpha;
    }
}

// err = |R|^2 (FSILS_NORMV squared again by CGRADV), the convergence test of
// the next loop top (CGRAD.f:205-215) by block 0, and
// P = (P + (errO/err) R) * (err/errO)  (OMPSUMV then OMPMULV, CGRAD.f:227-229).
// Block 0 may set the done flag while later blocks start; they then skip the
// P update, which is not used after the end.
__global__ __launch_bounds__(256) void k_cg_pupdate(size_t n, double* __restrict__ p,
                                                    const double* __restrict__ rr,
                                                    const double* __restrict__ part2, int m2,
                                                    CgState* __restrict__ st, int par)
{
    __shared__ double lds[4];
    __shared__ int done_s;
    if (threadIdx.x == 0) done_s = *(volatile int*)&st->done;
    __syncthreads();
    if (done_s) return;
    const double t = sum_partials(part2, m2, lds);
    const double nrm = sqrt(t);
    const double err = nrm * nrm;
    const double errO = st->e[par];
    const double r1 = errO / err;
    const double r2 = err / errO;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        p[q] = (p[q] + r1 * rr[q]) * r2;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->errO = errO;
        st->err = err;
        st->e[par ^ 1] = err;
        st->itr += 1;
        if (st->itr >= st->mItr) {
            st->done = 1;             // loop exhausted: suc stays false
        } else if (err < st->eps) {
            st->done = 1;
            st->suc = 1;
        }
    }
}

static unsigned vec_grid(size_t n)
{
    size_t g = (n + 255) / 256;
    if (g > 2048) g = 2048;
    if (g < 1) g = 1;
    return (unsigned)g;
}

PinnedPoll::~PinnedPoll()
{
    for (int q = 0; q < 2; ++q) {
        if (h[q]) (void)hipHostFree(h[q]);
        if (ev[q]) (void)hipEventDestroy(ev[q]);
    }
}

int PinnedPoll::ensure(size_t b)
{
    if (h[0] && b <= bytes) return SVFSI_OK;
    for (int q = 0; q < 2; ++q) {
        if (h[q]) (void)hipHostFree(h[q]);
        h[q] = nullptr;
        SV_HIP(hipHostMalloc(&h[q], b, hipHostMallocDefault));
        if (!ev[q]) SV_HIP(hipEventCreateWithFlags(&ev[q], hipEventDisableTiming));
    }
    bytes = b;
    return SVFSI_OK;
}

int PinnedPoll::post(int q, const void* dsrc, size_t n, hipStream_t s)
{
    SV_HIP(hipMemcpyAsync(h[q], dsrc, n, hipMemcpyDeviceToHost, s));
    SV_HIP(hipEventRecord(ev[q], s));
    return SVFSI_OK;
}

int PinnedPoll::wait(int q)
{
    SV_HIP(hipEventSynchronize(ev[q]));
    return SVFSI_OK;
}

GraphCache::~GraphCache()
{
    clear();
    if (cap) (void)hipStreamDestroy(cap);
}

void GraphCache::clear()
{
    for (auto& x : e) (void)hipGraphExecDestroy(x.exec);
    e.clear();
    sig.clear();
}

// Replays the graph stored under `key`, capturing `enqueue(stream)` into it
// first if needed.  Every buffer the captured launches use must be allocated
// before the capture (no allocation happens while a stream captures).
template <class F>
static int graph_run(GraphCache& gc, int key, hipStream_t s, F&& enqueue)
{
    for (auto& x : gc.e)
        if (x.key == key) {
            SV_HIP(hipGraphLaunch(x.exec, s));
            return SVFSI_OK;
        }
    if (!gc.cap) SV_HIP(hipStreamCreateWithFlags(&gc.cap, hipStreamNonBlocking));
    SV_HIP(hipStreamBeginCapture(gc.cap, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(gc.cap);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(gc.cap, &g);
    if (rc != SVFSI_OK || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rc != SVFSI_OK) return rc;
        SV_HIP(ec);
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    SV_HIP(ei);
    gc.e.push_back({key, ex});
    SV_HIP(hipGraphLaunch(ex, s));
    return SVFSI_OK;
}

// the coupled-face launches baked into a captured graph
static void cpl_sig(const Lhs& L, std::vector<uintptr_t>& sig)
{
    const CplSet& c = L.cpl;
    sig.insert(sig.end(), {(uintptr_t)c.active, (uintptr_t)c.gen, (uintptr_t)c.vm.p, (uintptr_t)c.coef.p,
                           (uintptr_t)c.glob.p, (uintptr_t)c.off.p, (uintptr_t)c.fdof.p, (uintptr_t)c.tmp.p});
}

// Graph replay is on unless SVFSI_GRAPHS=0; it is off while SpMV profiling
// is on (profiled launches carry events that a capture cannot hold).
static bool graphs_enabled()
{
    const char* v = getenv("SVFSI_GRAPHS");
    return !(v && v[0] == '0') && !prof().on;
}

int cg_solve(Lhs& L, FSILS_subLsType& ls, int dof, const double* K, double* R, hipStream_t s)
{
    if (dof < 2 || dof > 4) {
        set_error("CG: dof %d not supported", dof);
        return SVFSI_ERR_NOTSUP;
    }
    const size_t n = (size_t)L.nNo * dof;
    Work& w = L.ws;
    SV_TRY(w.X.ensure(n * sizeof(double)));
    SV_TRY(w.tmp.ensure(2 * n * sizeof(double)));
    SV_TRY(w.partial.ensure(kCgPartials * sizeof(double)));
    SV_TRY(w.state.ensure(kCgStateBytes));
    SV_TRY(w.poll.ensure(256));
    const CgBufs b{w.X.as<double>(), w.tmp.as<double>(), w.tmp.as<double>() + n, w.partial.as<double>(), w.state.p};
    return cg_core(L, ls, dof, K, R, s, b, graphs_enabled(), true);
}

int cg_core(Lhs& L, FSILS_subLsType& ls, int dof, const double* K, double* R, hipStream_t s, const CgBufs& b,
            bool useGraphs, bool pinned)
{
    static_assert(sizeof(CgState) <= kCgStateBytes, "CG state slot");
    if (dof < 2 || dof > 4) {
        set_error("CG: dof %d not supported", dof);
        return SVFSI_ERR_NOTSUP;
    }
    const size_t n = (size_t)L.nNo * dof;
    Work& w = L.ws;
    const bool vs = dof <= 3;  // dof 2 value streaming, dof 3 row lanes: 16 lanes per row either way
    const int rowsPerBlock = vs ? 256 / 16 : 4 * (kWave / (dof * dof));
    int gs = (L.nNo + rowsPerBlock - 1) / rowsPerBlock;
    // dof 3: the CSR-stream product when the pattern has row tiles
    int sT = 0, nT = 0;
    const int* tl = dof == 3 ? stream_tiles(L, 3, true, sT, nT) : nullptr;
    if (tl) gs = nT;
    if (gs > 1024) gs = 1024;
    if (gs < 1) gs = 1;
    unsigned gv = vec_grid(n);
    if (gv > 512) gv = 512;
    // partials: SpMV grid (<= 1024), vector grid (<= 512), launch_dot (512),
    // then the scalar result slot
    double* X = b.X;
    double* P = b.P;
    double* KP = b.KP;
    double* part1 = b.part;
    double* part2 = part1 + 1024;
    double* scal = part1 + kCgPartials - 8;
    CgState* st = static_cast<CgState*>(b.state);

    ls.callD = wall_seconds();
    ls.suc = 0;
    SV_TRY(launch_dot((int)n, R, R, part1, scal, s));
    double rr = 0.0;
    SV_HIP(hipMemcpyAsync(&rr, scal, sizeof(double), hipMemcpyDeviceToHost, s));
    SV_HIP(hipStreamSynchronize(s));
    ls.iNorm = std::sqrt(rr);
    double eps = std::fmax(ls.absTol, ls.relTol * ls.iNorm);
    eps = eps * eps;
    CgState h0{};
    h0.errO = ls.iNorm * ls.iNorm;
    h0.err = h0.errO;
    h0.e[0] = h0.err;
    h0.eps = eps;
    h0.mItr = ls.mItr;
    h0.itr = 0;
    h0.done = 0;
    h0.suc = 0;
    if (h0.err < eps) { h0.done = 1; h0.suc = 1; }
    if (ls.mItr <= 0) h0.done = 1;
    SV_HIP(hipMemcpyAsync(st, &h0, sizeof(CgState), hipMemcpyHostToDevice, s));
    SV_TRY(launch_fill(n, 0.0, X, s));
    SV_HIP(hipMemcpyAsync(P, R, n * sizeof(double), hipMemcpyDeviceToDevice, s));

    constexpr int chunk = 16; // even: iteration parity = position in the chunk
    // a chunk of iterations: the same launches every time, one graph
    auto iters = [&](hipStream_t cs) -> int {
        for (int it = 0; it < chunk; ++it) {
            const int par = it & 1;
            switch (dof) {
            case 2: hipLaunchKernelGGL((k_cg_spmv_dot_vs<2, 16>), dim3(gs), dim3(256), 0, cs, L.nNo, L.rowPtr.as<int>(), L.colPtr.as<int>(), K, (const double*)P, KP, part1, (const CgState*)st); break;
            case 3:
                if (tl && sT == 512)
                    hipLaunchKernelGGL(k_cg_sspmv_dot3<512>, dim3(gs), dim3(256), 0, cs, nT, tl, L.rowPtr.as<int>(), L.colPtr.as<int>(), K, (const double*)P, KP, part1, (const CgState*)st);
                else if (tl)
                    hipLaunchKernelGGL(k_cg_sspmv_dot3<256>, dim3(gs), dim3(256), 0, cs, nT, tl, L.rowPtr.as<int>(), L.colPtr.as<int>(), K, (const double*)P, KP, part1, (const CgState*)st);
                else
                    hipLaunchKernelGGL(k_cg_spmv_dot3r, dim3(gs), dim3(256), 0, cs, L.nNo, L.rowPtr.as<int>(), L.colPtr.as<int>(), K, (const double*)P, KP, part1, (const CgState*)st);
                break;
            default: hipLaunchKernelGGL(k_cg_spmv_dot<4>, dim3(gs), dim3(256), 0, cs, L.nNo, L.rowPtr.as<int>(), L.colPtr.as<int>(), K, (const double*)P, KP, part1, (const CgState*)st); break;
            }
            hipLaunchKernelGGL(k_cg_update, dim3(gv), dim3(256), 0, cs, n, X, R, (const double*)P, (const double*)KP,
                               (const double*)part1, gs, part2, st, par);
            hipLaunchKernelGGL(k_cg_pupdate, dim3(gv), dim3(256), 0, cs, n, P, (const double*)R,
                               (const double*)part2, (int)gv, st, par);
        }
        SV_HIP(hipGetLastError());
        return SVFSI_OK;
    };
    if (useGraphs) {
        const std::vector<uintptr_t> sig = {(uintptr_t)n, (uintptr_t)dof, (uintptr_t)K, (uintptr_t)R, (uintptr_t)X,
                                            (uintptr_t)P, (uintptr_t)part1, (uintptr_t)st, (uintptr_t)L.rowPtr.p,
                                            (uintptr_t)L.colPtr.p};
        if (sig != w.cgGraphs.sig) {
            w.cgGraphs.clear();
            w.cgGraphs.sig = sig;
        }
    }
    int posted = 0;
    CgState hs = h0;
    bool finished = h0.done != 0;
    int launched = 0;
    while (!finished) {
        const int q = posted & 1;
        if (useGraphs)
            SV_TRY(graph_run(w.cgGraphs, 0, s, iters));
        else
            SV_TRY(iters(s));
        launched += chunk;
        if (pinned) {
            SV_TRY(w.poll.post(q, st, sizeof(CgState), s));
            ++posted;
            // check the previous chunk while this one runs
            if (posted >= 2) {
                SV_TRY(w.poll.wait(q ^ 1));
                std::memcpy(&hs, w.poll.h[q ^ 1], sizeof(CgState));
                if (hs.done) finished = true;
            }
        } else {
            SV_HIP(hipMemcpyAsync(&hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
            SV_HIP(hipStreamSynchronize(s));
            if (hs.done) finished = true;
        }
        if (launched > ls.mItr + 2 * chunk) break; // safety: flag must be set by now
    }
    SV_HIP(hipStreamSynchronize(s));
    SV_HIP(hipMemcpyAsync(&hs, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    SV_HIP(hipMemcpyAsync(R, X, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    SV_HIP(hipStreamSynchronize(s));
    ls.suc = hs.suc;
    ls.itr = hs.itr;
    ls.cM = hs.itr;
    ls.fNorm = std::sqrt(hs.err);
    ls.callD = wall_seconds() - ls.callD;
    ls.dB = (hs.errO < 2.220446049250313e-16) ? 0.0 : 5.0 * std::log(hs.err / hs.errO);
    return SVFSI_OK;
}

// an integer tuning knob from the environment, read at each call
static int env_int(const char* k, int d)
{
    const char* e = getenv(k);
    return e ? atoi(e) : d;
}

// ---------------------------------------------------------------------------
// CGRAD_SCHUR (svFSILS/CGRAD.f:51-123): CG on S P = L P - Gt (G P), the
// pressure Schur complement of the NS solver (Gt = -G^T blockwise, so
// S = L + G^T G).  The same device-resident iteration as cg_solve (the
// update kernels are CGRADV's: CGRAD_SCHUR makes the same scalar and vector
// steps), with the operator in two row passes: GP = G P (k_cgs_pgp, which
// also makes the previous iteration's P update), then SP = L P - Gt GP and
// the partial sums of P . SP (k_cgs_sp).  Three launches per iteration.
// ---------------------------------------------------------------------------
// LANES lanes per row (a tet4 node has about 15 to 18 neighbours): lane l
// takes the row's blocks l, l+LANES, ..., so a
// wavefront's loads of G, Gt, L and colPtr are contiguous; the partial sums
// meet by shuffles.
template <int LANES>
__global__ __launch_bounds__(256) void k_cgs_sp(int nNo, const int* __restrict__ rowPtr,
                                                const int* __restrict__ colPtr, const double* __restrict__ Lm,
                                                const double* __restrict__ Gt, const double* __restrict__ P,
                                                const double* __restrict__ GP, double* __restrict__ SP,
                                                double* __restrict__ part, const CgState* __restrict__ st)
{
    if (st->done) return;
    __shared__ double lds[4];
    const int lane = threadIdx.x % LANES;
    const int rowsPerGrid = gridDim.x * (256 / LANES);
    double contrib = 0.0;
    for (int i = blockIdx.x * (256 / LANES) + threadIdx.x / LANES; i < nNo; i += rowsPerGrid) {
        double a = 0.0, d = 0.0;
        for (int j = rowPtr[i] + lane; j < rowPtr[i + 1]; j += LANES) {
            const int c = colPtr[j];
            a += Lm[j] * P[c];
            const double* k = Gt + (size_t)j * 3;
            const double* u = GP + (size_t)c * 3;
            d += k[0] * u[0] + k[1] * u[1] + k[2] * u[2];
        }
        for (int o = LANES / 2; o > 0; o >>= 1) {
            a += __shfl_xor(a, o, LANES);
            d += __shfl_xor(d, o, LANES);
        }
        if (lane == 0) {
            const double sp = a + (-1.0) * d;
            SP[i] = sp;
            contrib += P[i] * sp;
        }
    }
    const double sb = block_sum<256>(contrib, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = sb;
}

// The P update of the iteration just finished (k_cg_pupdate's step) fused
// with the next iteration's GP = G P: every block sums the |R|^2 partials
// itself, so each row can form the new P of its neighbours on the fly,
// (Pin + (errO/err) R) * (err/errO) with the same operations as the row
// that owns it; the owner writes it to Pout (a second buffer: other blocks
// still read Pin).  Block 0 takes over k_cg_pupdate's bookkeeping.  With
// first set (the solve's first iteration) P is Pin unchanged.
template <int LANES>
__global__ __launch_bounds__(256) void k_cgs_pgp(int nNo, const int* __restrict__ rowPtr,
                                                 const int* __restrict__ colPtr, const double* __restrict__ G,
                                                 const double* __restrict__ Pin, const double* __restrict__ R,
                                                 const double* __restrict__ part2, int m2, double* __restrict__ Pout,
                                                 double* __restrict__ GP, CgState* __restrict__ st, int par, int first)
{
    __shared__ double lds[4];
    __shared__ int done_s;
    if (threadIdx.x == 0) done_s = *(volatile int*)&st->done;
    __syncthreads();
    if (done_s) return;
    double r1 = 0.0, r2 = 1.0;
    if (!first) {
        const double t = sum_partials(part2, m2, lds);
        const double nrm = sqrt(t);
        const double err = nrm * nrm;
        const double errO = st->e[par ^ 1];
        r1 = errO / err;
        r2 = err / errO;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->errO = errO;
            st->err = err;
            st->e[par] = err;
            st->itr += 1;
            if (st->itr >= st->mItr) {
                st->done = 1;
            } else if (err < st->eps) {
                st->done = 1;
                st->suc = 1;
            }
        }
    }
    const int lane = threadIdx.x % LANES;
    const int rowsPerGrid = gridDim.x * (256 / LANES);
    for (int i = blockIdx.x * (256 / LANES) + threadIdx.x / LANES; i < nNo; i += rowsPerGrid) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int j = rowPtr[i] + lane; j < rowPtr[i + 1]; j += LANES) {
            const int c = colPtr[j];
            const double u = (Pin[c] + r1 * R[c]) * r2;
            a0 += G[(size_t)j * 3] * u;
            a1 += G[(size_t)j * 3 + 1] * u;
            a2 += G[(size_t)j * 3 + 2] * u;
        }
        for (int o = LANES / 2; o > 0; o >>= 1) {
            a0 += __shfl_xor(a0, o, LANES);
            a1 += __shfl_xor(a1, o, LANES);
            a2 += __shfl_xor(a2, o, LANES);
        }
        if (lane == 0) {
            Pout[i] = (Pin[i] + r1 * R[i]) * r2;
            GP[(size_t)i * 3] = a0;
            GP[(size_t)i * 3 + 1] = a1;
            GP[(size_t)i * 3 + 2] = a2;
        }
    }
}

// The same two row passes on the row tiles of the pattern (tiles of whole
// rows with at most T blocks, one 256-lane workgroup per tile, workgroups
// striding over the tiles): every lane loads T/256 of the tile's blocks with
// all loads issued before the first product, the per-block products go to
// LDS, and 8 lanes per row sum them with the strides and shuffles of the
// kernels above.  The rows' L P and Gt GP sums stay separate, as there.
constexpr int kCgsS = 8;

template <int T>
__global__ __launch_bounds__(256) void k_cgs_sp_t(int nTiles, const int* __restrict__ tile,
                                                  const int* __restrict__ rowPtr, const int* __restrict__ colPtr,
                                                  const double* __restrict__ Lm, const double* __restrict__ Gt,
                                                  const double* __restrict__ P, const double* __restrict__ GP,
                                                  double* __restrict__ SP, double* __restrict__ part,
                                                  const CgState* __restrict__ st)
{
    if (st->done) return;
    constexpr int U = T / 256;
    __shared__ double sA[T], sD[T];
    __shared__ int sR[T + 1];
    __shared__ double lds[4];
    const int tid = threadIdx.x;
    const int g = tid / kCgsS, l = tid % kCgsS;
    double contrib = 0.0;
    for (int b = blockIdx.x; b < nTiles; b += gridDim.x) {
        const int r0 = tile[b], r1 = tile[b + 1];
        const int k0 = rowPtr[r0];
        const int nb = rowPtr[r1] - k0, nr = r1 - r0;
        for (int i = tid; i <= nr; i += 256) sR[i] = rowPtr[r0 + i] - k0;
        int c[U];
        double lv[U], g0[U], g1[U], g2[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = u * 256 + tid;
            c[u] = j < nb ? colPtr[k0 + j] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = u * 256 + tid;
            if (j < nb) {
                const size_t k = (size_t)(k0 + j);
                lv[u] = Lm[k];
                g0[u] = Gt[k * 3];
                g1[u] = Gt[k * 3 + 1];
                g2[u] = Gt[k * 3 + 2];
            } else {
                lv[u] = g0[u] = g1[u] = g2[u] = 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = u * 256 + tid;
            const double pv = P[c[u]];
            const double* q = GP + (size_t)c[u] * 3;
            const double q0 = q[0], q1 = q[1], q2 = q[2];
            if (j < nb) {
                sA[j] = lv[u] * pv;
                sD[j] = g0[u] * q0 + g1[u] * q1 + g2[u] * q2;
            }
        }
        __syncthreads();
        for (int rr = g; rr < nr; rr += 256 / kCgsS) {
            double a = 0.0, d = 0.0;
            const int ke = sR[rr + 1];
            for (int j = sR[rr] + l; j < ke; j += kCgsS) {
                a += sA[j];
                d += sD[j];
            }
#pragma unroll
            for (int o = kCgsS / 2; o > 0; o >>= 1) {
                a += __shfl_xor(a, o, kWave);
                d += __shfl_xor(d, o, kWave);
            }
            if (l == 0) {
                const int i = r0 + rr;
                const double sp = a + (-1.0) * d;
                SP[i] = sp;
                contrib += P[i] * sp;
            }
        }
        __syncthreads();
    }
    const double sb = block_sum<256>(contrib, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = sb;
}

template <int T>
__global__ __launch_bounds__(256) void k_cgs_pgp_t(int nTiles, const int* __restrict__ tile,
                                                   const int* __restrict__ rowPtr, const int* __restrict__ colPtr,
                                                   const double* __restrict__ G, const double* __restrict__ Pin,
                                                   const double* __restrict__ R, const double* __restrict__ part2,
                                                   int m2, double* __restrict__ Pout, double* __restrict__ GP,
                                                   CgState* __restrict__ st, int par, int first)
{
    constexpr int U = T / 256;
    __shared__ double s0[T], s1[T], s2[T];
    __shared__ int sR[T + 1];
    __shared__ double lds[4];
    __shared__ int done_s;
    if (threadIdx.x == 0) done_s = *(volatile int*)&st->done;
    __syncthreads();
    if (done_s) return;
    double r1 = 0.0, r2 = 1.0;
    if (!first) {
        const double t = sum_partials(part2, m2, lds);
        const double nrm = sqrt(t);
        const double err = nrm * nrm;
        const double errO = st->e[par ^ 1];
        r1 = errO / err;
        r2 = err / errO;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->errO = errO;
            st->err = err;
            st->e[par] = err;
            st->itr += 1;
            if (st->itr >= st->mItr) {
                st->done = 1;
            } else if (err < st->eps) {
                st->done = 1;
                st->suc = 1;
            }
        }
    }
    const int tid = threadIdx.x;
    const int g = tid / kCgsS, l = tid % kCgsS;
    for (int b = blockIdx.x; b < nTiles; b += gridDim.x) {
        const int r0 = tile[b], r1t = tile[b + 1];
        const int k0 = rowPtr[r0];
        const int nb = rowPtr[r1t] - k0, nr = r1t - r0;
        for (int i = tid; i <= nr; i += 256) sR[i] = rowPtr[r0 + i] - k0;
        int c[U];
        double g0[U], g1[U], g2[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = u * 256 + tid;
            c[u] = j < nb ? colPtr[k0 + j] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = u * 256 + tid;
            if (j < nb) {
                const size_t k = (size_t)(k0 + j);
                g0[u] = G[k * 3];
                g1[u] = G[k * 3 + 1];
                g2[u] = G[k * 3 + 2];
            } else {
                g0[u] = g1[u] = g2[u] = 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = u * 256 + tid;
            const double uv = (Pin[c[u]] + r1 * R[c[u]]) * r2;
            if (j < nb) {
                s0[j] = g0[u] * uv;
                s1[j] = g1[u] * uv;
                s2[j] = g2[u] * uv;
            }
        }
        __syncthreads();
        for (int rr = g; rr < nr; rr += 256 / kCgsS) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            const int ke = sR[rr + 1];
            for (int j = sR[rr] + l; j < ke; j += kCgsS) {
                a0 += s0[j];
                a1 += s1[j];
                a2 += s2[j];
            }
#pragma unroll
            for (int o = kCgsS / 2; o > 0; o >>= 1) {
                a0 += __shfl_xor(a0, o, kWave);
                a1 += __shfl_xor(a1, o, kWave);
                a2 += __shfl_xor(a2, o, kWave);
            }
            if (l == 0) {
                const int i = r0 + rr;
                Pout[i] = (Pin[i] + r1 * R[i]) * r2;
                GP[(size_t)i * 3] = a0;
                GP[(size_t)i * 3 + 1] = a1;
                GP[(size_t)i * 3 + 2] = a2;
            }
        }
        __syncthreads();
    }
}

int cg_schur_solve(Lhs& L, FSILS_subLsType& ls, const double* Gt, const double* G, const double* Lm, double* R,
                   double* X, double* P, double* P2, double* SP, double* GP, hipStream_t s)
{
    const int nNo = L.nNo;
    const size_t n = (size_t)nNo;
    Work& w = L.ws;
    // lanes per row (tuning knob SVFSI_CGS_LANES = 8, 16 or 32; 8 measured
    // fastest on the 200k-tet cylinder, 32 slowest)
    static const int lanes = [] {
        const char* e = getenv("SVFSI_CGS_LANES");
        const int v = e ? atoi(e) : 8;
        return (v == 16 || v == 32) ? v : 8;
    }();
    // Row tiles of 512 blocks from 2M blocks on, the row-per-lane-group
    // kernels below: ms per NS step (3.2M-tet cylinder / 200k-tet one)
    // 131.9 / 8.35 with 512-block tiles, 144.6 / 8.55 with 1024, 154.9 /
    // 8.27 without; on the small matrix the tile kernels' longer dependency
    // chain (tile bounds, LDS pass, barriers) outweighs their loads in
    // flight.  Streaming (non-temporal) loads of L, Gt, G measured slower
    // (104 against 92 us per k_cgs_sp_t on the large cylinder), and so did
    // an XCD-contiguous tile order (94 us).  SVFSI_CGS_TILE = 0, 512 or 1024
    // forces the choice (read at each call).
    const int cgsEnv = [] {
        const char* e = getenv("SVFSI_CGS_TILE");
        if (!e) return -1;
        const int v = atoi(e);
        return (v == 512 || v == 1024) ? v : 0;
    }();
    const int cgsTile = cgsEnv >= 0 ? cgsEnv : ((size_t)L.nnz >= ((size_t)2 << 20) ? 512 : 0);
    const int ti = cgsTile == 1024 ? 2 : 1;
    const int* tl = (cgsTile && L.nTiles[ti]) ? L.tiles[ti].as<int>() : nullptr;
    const int nT = tl ? L.nTiles[ti] : 0;
    constexpr int kMaxGs = 2048;
    int gs = tl ? nT : (nNo + 256 / lanes - 1) / (256 / lanes);
    if (gs > kMaxGs) gs = kMaxGs;
    if (gs < 1) gs = 1;
    unsigned gv = vec_grid(n);
    if (gv > 512) gv = 512;
    const size_t nPart = kMaxGs + 512 + 16;
    SV_TRY(w.partial.ensure(nPart * sizeof(double)));
    SV_TRY(w.state.ensure(sizeof(CgState)));
    double* part1 = w.partial.as<double>();
    double* part2 = part1 + kMaxGs;
    double* scal = part1 + nPart - 8;
    CgState* st = w.state.as<CgState>();

    const double time = wall_seconds();
    ls.suc = 0;
    SV_TRY(launch_dot((int)n, R, R, part1, scal, s));
    double rr = 0.0;
    SV_HIP(hipMemcpyAsync(&rr, scal, sizeof(double), hipMemcpyDeviceToHost, s));
    SV_HIP(hipStreamSynchronize(s));
    ls.iNorm = std::sqrt(rr);
    double eps = std::fmax(ls.absTol, ls.relTol * ls.iNorm);
    eps = eps * eps;
    CgState h0{};
    h0.errO = ls.iNorm * ls.iNorm;
    h0.err = h0.errO;
    h0.e[0] = h0.err;
    h0.eps = eps;
    h0.mItr = ls.mItr;
    if (h0.err < eps) { h0.done = 1; h0.suc = 1; }
    if (ls.mItr <= 0) h0.done = 1;
    SV_HIP(hipMemcpyAsync(st, &h0, sizeof(CgState), hipMemcpyHostToDevice, s));
    SV_HIP(hipMemsetAsync(X, 0, n * sizeof(double), s));
    // P double-buffered: the fused update reads one buffer and writes the
    // other; chunk position it reads buf[(it+1)&1] and writes buf[it&1]
    double* buf[2] = {P, P2};
    SV_HIP(hipMemcpyAsync(buf[1], R, n * sizeof(double), hipMemcpyDeviceToDevice, s));

    PinnedPoll& poll = w.poll;
    SV_TRY(poll.ensure(256));
    constexpr int chunk = 16; // even: iteration parity = position in the chunk
    const bool cpl = L.cpl.active;
    const int* rp = L.rowPtr.as<int>();
    const int* cp = L.colPtr.as<int>();
    // a chunk of iterations (the first chunk of a solve differs only in its
    // first position, which has no finished iteration to update P for)
    auto iters = [&](hipStream_t cs, bool firstChunk) -> int {
        for (int it = 0; it < chunk; ++it) {
            const int par = it & 1;
            const double* Pin = buf[(it + 1) & 1];
            double* Pc = buf[it & 1];
            const int first = (firstChunk && it == 0) ? 1 : 0;
            if (tl && cgsTile == 1024)
                hipLaunchKernelGGL(k_cgs_pgp_t<1024>, dim3(gs), dim3(256), 0, cs, nT, tl, rp, cp, G, Pin,
                                   (const double*)R, (const double*)part2, (int)gv, Pc, GP, st, par, first);
            else if (tl)
                hipLaunchKernelGGL(k_cgs_pgp_t<512>, dim3(gs), dim3(256), 0, cs, nT, tl, rp, cp, G, Pin,
                                   (const double*)R, (const double*)part2, (int)gv, Pc, GP, st, par, first);
            else if (lanes == 8)
                hipLaunchKernelGGL(k_cgs_pgp<8>, dim3(gs), dim3(256), 0, cs, nNo, rp, cp, G, Pin, (const double*)R,
                                   (const double*)part2, (int)gv, Pc, GP, st, par, first);
            else if (lanes == 32)
                hipLaunchKernelGGL(k_cgs_pgp<32>, dim3(gs), dim3(256), 0, cs, nNo, rp, cp, G, Pin, (const double*)R,
                                   (const double*)part2, (int)gv, Pc, GP, st, par, first);
            else
                hipLaunchKernelGGL(k_cgs_pgp<16>, dim3(gs), dim3(256), 0, cs, nNo, rp, cp, G, Pin, (const double*)R,
                                   (const double*)part2, (int)gv, Pc, GP, st, par, first);
            // coupled faces: GP <- PRE GP (CGRAD.f:89-92)
            if (cpl) SV_TRY(cpl_apply(L, 1, 3, GP, GP, cs, &st->done));
            {
                // algorithmic bytes: L, Gt, colPtr, rowPtr; P and GP once; SP
                ProfScope ps(4, 36.0 * L.nnz + 4.0 * (nNo + 1) + 40.0 * nNo);
                if (tl && cgsTile == 1024)
                    SV_LAUNCH(ps, k_cgs_sp_t<1024>, dim3(gs), dim3(256), 0, cs, nT, tl, rp, cp, Lm, Gt,
                              (const double*)Pc, (const double*)GP, SP, part1, (const CgState*)st);
                else if (tl)
                    SV_LAUNCH(ps, k_cgs_sp_t<512>, dim3(gs), dim3(256), 0, cs, nT, tl, rp, cp, Lm, Gt,
                              (const double*)Pc, (const double*)GP, SP, part1, (const CgState*)st);
                else if (lanes == 8)
                    SV_LAUNCH(ps, k_cgs_sp<8>, dim3(gs), dim3(256), 0, cs, nNo, rp, cp, Lm, Gt, (const double*)Pc,
                              (const double*)GP, SP, part1, (const CgState*)st);
                else if (lanes == 32)
                    SV_LAUNCH(ps, k_cgs_sp<32>, dim3(gs), dim3(256), 0, cs, nNo, rp, cp, Lm, Gt, (const double*)Pc,
                              (const double*)GP, SP, part1, (const CgState*)st);
                else
                    SV_LAUNCH(ps, k_cgs_sp<16>, dim3(gs), dim3(256), 0, cs, nNo, rp, cp, Lm, Gt, (const double*)Pc,
                              (const double*)GP, SP, part1, (const CgState*)st);
            }
            hipLaunchKernelGGL(k_cg_update, dim3(gv), dim3(256), 0, cs, n, X, R, (const double*)Pc, (const double*)SP,
                               (const double*)part1, gs, part2, st, par);
        }
        SV_HIP(hipGetLastError());
        return SVFSI_OK;
    };
    const bool useGraphs = graphs_enabled();
    if (useGraphs) {
        std::vector<uintptr_t> sig = {(uintptr_t)n, (uintptr_t)Gt, (uintptr_t)G, (uintptr_t)Lm, (uintptr_t)R,
                                      (uintptr_t)X, (uintptr_t)P, (uintptr_t)P2, (uintptr_t)SP, (uintptr_t)GP,
                                      (uintptr_t)part1, (uintptr_t)st, (uintptr_t)rp, (uintptr_t)cp, (uintptr_t)cpl,
                                      (uintptr_t)lanes, (uintptr_t)tl};
        if (cpl) cpl_sig(L, sig);
        if (sig != w.csGraphs.sig) {
            w.csGraphs.clear();
            w.csGraphs.sig = sig;
        }
    }
    int posted = 0;
    CgState hs = h0;
    bool finished = h0.done != 0;
    int launched = 0;
    while (!finished) {
        const int q = posted & 1;
        const bool firstChunk = launched == 0;
        if (useGraphs)
            SV_TRY(graph_run(w.csGraphs, firstChunk ? 0 : 1, s, [&](hipStream_t cs) { return iters(cs, firstChunk); }));
        else
            SV_TRY(iters(s, firstChunk));
        launched += chunk;
        SV_TRY(poll.post(q, st, sizeof(CgState), s));
        ++posted;
        if (posted >= 2) {
            SV_TRY(poll.wait(q ^ 1));
            std::memcpy(&hs, poll.h[q ^ 1], sizeof(CgState));
            if (hs.done) finished = true;
        }
        // the last iteration's bookkeeping runs at the next chunk's start:
        // one chunk beyond mItr at most
        if (launched > ls.mItr + 3 * chunk) break; // safety: the flag is set by now
    }
    SV_HIP(hipStreamSynchronize(s));
    SV_HIP(hipMemcpy(&hs, st, sizeof(CgState), hipMemcpyDeviceToHost));
    SV_HIP(hipMemcpyAsync(R, X, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    SV_HIP(hipStreamSynchronize(s));
    ls.suc = hs.suc;
    ls.fNorm = std::sqrt(hs.err);
    ls.callD = wall_seconds() - time + ls.callD;
    ls.itr = ls.itr + hs.itr;
    ls.dB = (hs.errO < 2.220446049250313e-16) ? 0.0 : 5.0 * std::log(hs.err / hs.errO);
    return SVFSI_OK;
}

// ===========================================================================
// GMRESV (svFSILS/GMRES.f:273-431), templated on the basis storage type BT
// (double for the FP64 path; float for the inner solver of the mixed path,
// whose dot products and Hessenberg arithmetic stay FP64).
// ===========================================================================
struct GmState {
    double eps;
    double fNorm;
    int stop, suc, iconv, iters;  // iters: Arnoldi steps run (accumulates)
    unsigned int counter;
    int pad2;
};

struct GmArrays {
    double* H;     // (sD+1, sD) column-major: H(j,i) at (i-1)*(sD+1)+(j-1)
    double* c;     // sD
    double* sn;    // sD
    double* err;   // sD+1
    double* y;     // sD
    double* hraw;  // sD+2: CGS coefficients of the current step, then 1/h(i+1,i)
    double* Hu;    // (sD+1, sD): the unrotated Arnoldi Hessenberg (inner cycle)
};

static size_t gm_arrays_doubles(int sD)
{
    return 2 * (size_t)(sD + 1) * sD + 3 * sD + (sD + 1) + (sD + 3);
}

static GmArrays gm_carve(double* base, int sD)
{
    GmArrays a;
    a.H = base;
    a.c = a.H + (size_t)(sD + 1) * sD;
    a.sn = a.c + sD;
    a.err = a.sn + sD;
    a.y = a.err + (sD + 1);
    a.hraw = a.y + sD;
    a.Hu = a.hraw + (sD + 3);
    return a;
}

// Initial state of a GMRES solve.
__global__ void k_gm_init(GmState* __restrict__ st, double eps)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    st->eps = eps;
    st->fNorm = 0.0;
    st->stop = 0;
    st->suc = 0;
    st->iconv = 0;
    st->iters = 0;
    st->counter = 0u;
}

// u0 = R - u0 and |u0|; last block: err(1) = |u0|, cycle state reset.
template <typename BT>
__global__ __launch_bounds__(256) void k_gm_resid(size_t n, const BT* __restrict__ R,
                                                  BT* __restrict__ u0,
                                                  double* __restrict__ partial,
                                                  GmState* __restrict__ st, GmArrays A)
{
    __shared__ double lds[4];
    __shared__ int flag;
    double s = 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        const BT v = R[q] - u0[q];
        u0[q] = v;
        s += (double)v * (double)v;
    }
    s = block_sum<256>(s, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
    if (last_block_done(&st->counter, &flag)) {
        double t = 0.0;
        for (int q = threadIdx.x; q < (int)gridDim.x; q += 256) t += load_agent(partial + q);
        t = block_sum<256>(t, lds);
        if (threadIdx.x == 0) {
            A.err[0] = sqrt(t);
            st->stop = 0;
            st->iconv = 0;
        }
    }
}

// u0 = u0 / err(1)   (GMRES.f:336: a division, not a reciprocal product)
template <typename BT>
__global__ __launch_bounds__(256) void k_gm_div(size_t n, BT* __restrict__ u0, GmArrays A)
{
    const double e0 = A.err[0];
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        u0[q] = (BT)((double)u0[q] / e0);
}

constexpr int kMaxSD = 256;

// Vector-load helpers: E consecutive elements as one 16-byte load.
template <typename BT> struct VecE;
template <> struct VecE<double> {
    static constexpr int E = 2;
    typedef double v __attribute__((ext_vector_type(2)));
};
template <> struct VecE<float> {
    static constexpr int E = 4;
    typedef float v __attribute__((ext_vector_type(4)));
};

template <typename BT>
__device__ __forceinline__ void loadE(const BT* __restrict__ p, size_t q, size_t q1, double* o)
{
    constexpr int E = VecE<BT>::E;
    if (q + E <= q1 && ((q & (E - 1)) == 0)) {
        const typename VecE<BT>::v t = *reinterpret_cast<const typename VecE<BT>::v*>(p + q);
#pragma unroll
        for (int e = 0; e < E; ++e) o[e] = (double)t[e];
    } else {
#pragma unroll
        for (int e = 0; e < E; ++e) o[e] = (q + e < q1) ? (double)p[q + e] : 0.0;
    }
}

// h(j,i) = <u_j, u_{i+1}> for j = 1..i+1 (GMRES.f:337-340), per-block
// partial sums.  i is the 1-based Arnoldi step; u_j is basis + (j-1)*n.
// Grid: (row chunks) x (groups of 8 coefficients).  A block keeps its chunk
// of the new vector in registers (E elements per lane, 16-byte loads) and
// forms eight dot products at once, so their loads and lane-shuffle
// reductions interleave; the four wavefront sums meet in LDS.  There is no
// intra-kernel cross-block synchronisation (on a multi-XCD part an
// agent-scope fence writes back the XCD's L2): the partials are summed by
// k_gm_givens.
template <typename BT>
__global__ __launch_bounds__(256) void k_gm_mdot(size_t n, int i, const BT* __restrict__ basis,
                                                 double* __restrict__ partial,
                                                 const GmState* __restrict__ st)
{
    if (st->stop) return;
    constexpr int E = VecE<BT>::E;
    __shared__ double red[4][8];
    const int nbr = gridDim.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t chunk = (size_t)256 * E;
    const size_t q0 = (size_t)blockIdx.x * chunk;
    const size_t q1 = (q0 + chunk < n) ? q0 + chunk : n;
    const size_t q = q0 + (size_t)threadIdx.x * E;
    const int j0 = blockIdx.y * 8;
    double wv[E];
    loadE(basis + (size_t)i * n, q, q1, wv);
    double s[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        s[b] = 0.0;
        const int jj = j0 + b;
        if (jj <= i) {
            double uv[E];
            loadE(basis + (size_t)jj * n, q, q1, uv);
#pragma unroll
            for (int e = 0; e < E; ++e) s[b] += uv[e] * wv[e];
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int b = 0; b < 8; ++b) s[b] += __shfl_xor(s[b], o, kWave);
    if (lane == 0)
#pragma unroll
        for (int b = 0; b < 8; ++b) red[wave][b] = s[b];
    __syncthreads();
    if (threadIdx.x < 8) {
        const int jj = j0 + threadIdx.x;
        if (jj <= i)
            partial[(size_t)jj * nbr + blockIdx.x] =
                ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
    }
}

// Many row chunks (large systems): one workgroup per coefficient jj sums
// the nb partials of k_gm_mdot in a fixed order and leaves the sum in the
// coefficient's first slot, so that k_gm_givens reads one value per
// coefficient instead of summing (i+1) nb partials in a single workgroup
// (about 100 us per Arnoldi step at 3.5k chunks).
__global__ __launch_bounds__(256) void k_gm_fold(int nb, double* __restrict__ partial,
                                                 const GmState* __restrict__ st)
{
    if (st->stop) return;
    __shared__ double lds[4];
    double* pp = partial + (size_t)blockIdx.x * nb;
    double t0 = 0.0, t1 = 0.0;
    int b = threadIdx.x;
    for (; b + 256 < nb; b += 512) {
        t0 += pp[b];
        t1 += pp[b + 256];
    }
    if (b < nb) t0 += pp[b];
    const double sum = block_sum<256>(t0 + t1, lds);
    if (threadIdx.x == 0) pp[0] = sum;
}

// Row chunks above which the multi-dot partials are folded first
// (SVFSI_FOLD_CHUNKS overrides, read at each solve).
constexpr int kFoldChunks = 1024;

// The serial part of an Arnoldi step (one thread): h(i+1,i) (Pythagorean
// for passes 0-2, hsh[i] = |w|^2 for pass 3), the previous rotations, the
// new Givens rotation, the Hessenberg column, the residual estimate and the
// convergence test (GMRES.f:342-366).  hsh: the step's coefficients;
// c, sn: the previous rotations.
__device__ void givens_tail(int i, int sD, int pass, double* hsh, const double* c, const double* sn,
                            GmState* __restrict__ st, GmArrays A)
{
    st->iters += 1;
    double hii = hsh[i];
    if (pass != 3)
        for (int jj = 0; jj < i; ++jj) hii = hii - hsh[jj] * hsh[jj];
    hii = sqrt(fabs(hii));
    if (pass == 2)
        for (int jj = 0; jj < i; ++jj) hsh[jj] = A.y[jj] + hsh[jj];
    hsh[i] = hii;
    A.hraw[sD + 2] = 1.0 / hii;  // OMPMULV(1/h(i+1,i)) factor for k_gm_update
    for (int jj = 1; jj <= i - 1; ++jj) {
        const double tmp = c[jj - 1] * hsh[jj - 1] + sn[jj - 1] * hsh[jj];
        hsh[jj] = -sn[jj - 1] * hsh[jj - 1] + c[jj - 1] * hsh[jj];
        hsh[jj - 1] = tmp;
    }
    const double tmp = sqrt(hsh[i - 1] * hsh[i - 1] + hsh[i] * hsh[i]);
    const double ci = hsh[i - 1] / tmp;
    const double si = hsh[i] / tmp;
    hsh[i - 1] = tmp;
    hsh[i] = 0.0;
    A.c[i - 1] = ci;
    A.sn[i - 1] = si;
    double* Hc = A.H + (size_t)(i - 1) * (sD + 1); // column i: Hc[j-1] = h(j,i)
    for (int jj = 0; jj <= i; ++jj) Hc[jj] = hsh[jj];
    const double ei = A.err[i - 1];
    const double eNew = -si * ei;
    A.err[i] = eNew;
    A.err[i - 1] = ci * ei;
    if (fabs(eNew) < st->eps) {
        st->suc = 1;
        st->stop = 1;
        st->iconv = i;
    } else if (i >= sD) {
        st->stop = 1;
        st->iconv = sD;
    }
}

// One block: sums the per-block partials in block order (deterministic),
// then the Pythagorean h(i+1,i) (GMRES.f:342-347), the previous rotations,
// the new Givens rotation, the residual estimate and the convergence test
// (351-366), all in LDS.
// pass 0: single classical Gram-Schmidt pass with the Pythagorean norm
// (GMRESV).  pass 1 / 2: the two passes of CGS2 (used by the flexible outer
// solver): pass 1 only stores the first coefficients (in y, free during a
// cycle) for k_gm_update; pass 2 takes h(i+1,i) from the twice-orthogonalised
// vector and adds the two coefficient sets before the Givens rotations.
// pass 3: a single classical pass with the explicit norm (after pass 1 and
// k_gm_update with its |w|^2 partials): h(i+1,i) = |w|.
// Coefficient jj's nb partials are at partial[jj * ldp ...]: ldp = nb for
// the partials of k_gm_mdot itself, nb = 1 once k_gm_fold has summed them.
__global__ __launch_bounds__(256) void k_gm_givens(int i, int sD, int nb, int ldp, const double* __restrict__ partial,
                                                   GmState* __restrict__ st, GmArrays A, int pass)
{
    if (st->stop) return;
    __shared__ double hsh[kMaxSD + 2];
    __shared__ double cs[2][kMaxSD];
    __shared__ double lds[4];
    if (pass == 3) {
        // single pass with the explicit norm: partial[0..nb) hold the
        // partial sums of |w|^2 after k_gm_update, the coefficients are the
        // pass-1 ones kept in y
        double t = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) t += partial[b];
        t = block_sum<256>(t, lds);
        for (int jj = threadIdx.x; jj < i; jj += 256) hsh[jj] = A.y[jj];
        if (threadIdx.x == 0) hsh[i] = t;
    }
    // T lanes per coefficient (power of two <= 64), fixed-order sums
    int npow = 1;
    while (npow < i + 1) npow <<= 1;
    int T = 256 / npow;
    if (T > 64) T = 64;
    if (T < 1) T = 1;
    const int per = 256 / T;
    for (int jj0 = 0; jj0 <= i && pass != 3; jj0 += per) {
        const int jj = jj0 + threadIdx.x / T;
        const int b0 = threadIdx.x % T;
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
        if (jj <= i) {
            const double* pp = partial + (size_t)jj * ldp;
            int b = b0;
            for (; b + 3 * T < nb; b += 4 * T) {
                t0 += pp[b];
                t1 += pp[b + T];
                t2 += pp[b + 2 * T];
                t3 += pp[b + 3 * T];
            }
            for (; b < nb; b += T) t0 += pp[b];
        }
        double t = (t0 + t1) + (t2 + t3);
        for (int o = T >> 1; o > 0; o >>= 1) t += __shfl_xor(t, o, kWave);
        if (jj <= i && b0 == 0) hsh[jj] = t;
    }
    for (int jj = threadIdx.x; jj < i - 1; jj += 256) {
        cs[0][jj] = A.c[jj];
        cs[1][jj] = A.sn[jj];
    }
    __syncthreads();
    // CGS coefficients for k_gm_update (pre-rotation values)
    for (int jj = threadIdx.x; jj <= i; jj += 256) A.hraw[jj] = hsh[jj];
    if (pass == 1) {
        for (int jj = threadIdx.x; jj < i; jj += 256) A.y[jj] = hsh[jj];
        if (threadIdx.x == 0) A.hraw[sD + 2] = 1.0;
        return;
    }
    if (threadIdx.x != 0) return;
    givens_tail(i, sD, pass, hsh, cs[0], cs[1], st, A);
}

// u_{i+1} = (u_{i+1} - sum_j h(j,i) u_j) * (1/h(i+1,i))   (GMRES.f:342-349)
// The pre-rotation coefficients were saved in hraw by k_gm_givens.  Each
// lane updates E consecutive elements with 16-byte loads; the sweep over j
// keeps the reference's order per element.
// With normPart given, each block also leaves the partial sum of the
// updated vector's squares in normPart[blockIdx.x].
template <typename BT>
__global__ __launch_bounds__(256) void k_gm_update(size_t n, int i, int sD, BT* __restrict__ basis,
                                                   const GmState* __restrict__ st, GmArrays A,
                                                   double* __restrict__ normPart)
{
    if (st->stop) return;
    constexpr int E = VecE<BT>::E;
    __shared__ double hc[kMaxSD + 1];
    __shared__ double lds[4];
    double nrm = 0.0;
    for (int jj = threadIdx.x; jj < i; jj += 256) hc[jj] = -A.hraw[jj];
    __syncthreads();
    const double inv = A.hraw[sD + 2];
    BT* w = basis + (size_t)i * n;
    for (size_t q = ((size_t)blockIdx.x * 256 + threadIdx.x) * E; q < n; q += (size_t)gridDim.x * 256 * E) {
        double v[E];
        loadE(w, q, n, v);
        int jj = 0;
        for (; jj + 4 <= i; jj += 4) {   // four loads in flight, same order
            double u0[E], u1[E], u2[E], u3[E];
            loadE(basis + (size_t)jj * n, q, n, u0);
            loadE(basis + (size_t)(jj + 1) * n, q, n, u1);
            loadE(basis + (size_t)(jj + 2) * n, q, n, u2);
            loadE(basis + (size_t)(jj + 3) * n, q, n, u3);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                v[e] = v[e] + hc[jj] * u0[e];
                v[e] = v[e] + hc[jj + 1] * u1[e];
                v[e] = v[e] + hc[jj + 2] * u2[e];
                v[e] = v[e] + hc[jj + 3] * u3[e];
            }
        }
        for (; jj < i; ++jj) {
            double u[E];
            loadE(basis + (size_t)jj * n, q, n, u);
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = v[e] + hc[jj] * u[e];
        }
#pragma unroll
        for (int e = 0; e < E; ++e)
            if (q + e < n) {
                const double o = v[e] * inv;
                w[q + e] = (BT)o;
                nrm += o * o;
            }
    }
    if (normPart) {
        const double sb = block_sum<256>(nrm, lds);
        if (threadIdx.x == 0) normPart[blockIdx.x] = sb;
    }
}

// The explicit-norm pass's end in one launch: every workgroup sums the nb
// partials of |w|^2 (the order of k_gm_givens pass 3) and scales its rows
// of w by 1/|w|; workgroup 0 also makes the rotation and the convergence
// test.  Once that has set stop the step's result is not used, so
// workgroups that start later and skip their rows change nothing.
template <typename BT>
__global__ __launch_bounds__(256) void k_gm_normscale(size_t n, int i, int sD, int nb,
                                                      const double* __restrict__ partial, BT* __restrict__ basis,
                                                      GmState* __restrict__ st, GmArrays A)
{
    if (st->stop) return;
    __shared__ double lds[4];
    __shared__ double hsh[kMaxSD + 2];
    __shared__ double cs[2][kMaxSD];
    double t = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) t += partial[b];
    t = block_sum<256>(t, lds);
    const double inv = 1.0 / sqrt(fabs(t));
    BT* w = basis + (size_t)i * n;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        w[q] = (BT)((double)w[q] * inv);
    if (blockIdx.x != 0) return;
    for (int jj = threadIdx.x; jj < i; jj += 256) hsh[jj] = A.y[jj];
    for (int jj = threadIdx.x; jj < i - 1; jj += 256) {
        cs[0][jj] = A.c[jj];
        cs[1][jj] = A.sn[jj];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    hsh[i] = t;
    givens_tail(i, sD, 3, hsh, cs[0], cs[1], st, A);
}

// w *= 1/h(i+1,i) (the explicit-norm single pass)
template <typename BT>
__global__ __launch_bounds__(256) void k_gm_scale(size_t n, int i, int sD, BT* __restrict__ basis,
                                                  const GmState* __restrict__ st, GmArrays A)
{
    if (st->stop) return;
    const double inv = A.hraw[sD + 2];
    BT* w = basis + (size_t)i * n;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        w[q] = (BT)((double)w[q] * inv);
}

// Step end of the device-driven inner cycle, fused: every block sums the
// multi-dot partials (the fixed order of k_gm_givens), forms h(i+1,i) by the
// Pythagorean rule and updates its rows of u_{i+1}
// (u_{i+1} - sum_j h(j,i) u_j) / h(i+1,i); block 0 then applies the Givens
// rotations and the convergence test.  Once block 0 has set the stop flag a
// block that starts later skips its rows: u_{i+1} is not used after a stop.
// PRE: the coefficients were already summed into A.hraw by k_gm_hsum (the
// fused SpMV + dot path); otherwise every block sums the multi-dot partials.
template <typename BT, bool PRE = false>
__global__ __launch_bounds__(256) void k_gm_rotupdate(size_t n, int i, int sD, int nb,
                                                      const double* __restrict__ partial, BT* __restrict__ basis,
                                                      GmState* __restrict__ st, GmArrays A)
{
    // block 0 may set the stop flag while other blocks start: one read per
    // block keeps the exit uniform across the block's barriers
    __shared__ int stop_s;
    if (threadIdx.x == 0) stop_s = *(volatile int*)&st->stop;
    __syncthreads();
    if (stop_s) return;
    __shared__ double hsh[kMaxSD + 2];
    __shared__ double hc[kMaxSD + 1];
    __shared__ double inv_s;
    if constexpr (PRE) {
        for (int jj = threadIdx.x; jj <= i; jj += 256) hsh[jj] = A.hraw[jj];
    } else {
        int npow = 1;
        while (npow < i + 1) npow <<= 1;
        int T = 256 / npow;
        if (T > 64) T = 64;
        if (T < 1) T = 1;
        const int per = 256 / T;
        for (int jj0 = 0; jj0 <= i; jj0 += per) {
            const int jj = jj0 + threadIdx.x / T;
            const int b0 = threadIdx.x % T;
            double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
            if (jj <= i) {
                const double* pp = partial + (size_t)jj * nb;
                int b = b0;
                for (; b + 3 * T < nb; b += 4 * T) {
                    t0 += pp[b];
                    t1 += pp[b + T];
                    t2 += pp[b + 2 * T];
                    t3 += pp[b + 3 * T];
                }
                for (; b < nb; b += T) t0 += pp[b];
            }
            double t = (t0 + t1) + (t2 + t3);
            for (int o = T >> 1; o > 0; o >>= 1) t += __shfl_xor(t, o, kWave);
            if (jj <= i && b0 == 0) hsh[jj] = t;
        }
    }
    __syncthreads();
    // Pythagorean h(i+1,i) = sqrt|<w,w> - sum h_j^2|: the sum over j < i by
    // the first wavefront (fixed order), so no thread walks the column alone
    __shared__ double cs_s[2][kMaxSD];
    for (int jj = threadIdx.x; jj < i; jj += 256) hc[jj] = -hsh[jj];
    if (blockIdx.x == 0)
        for (int jj = threadIdx.x; jj < i - 1; jj += 256) {
            cs_s[0][jj] = A.c[jj];
            cs_s[1][jj] = A.sn[jj];
        }
    if (threadIdx.x < kWave) {
        double q = 0.0;
        for (int jj = threadIdx.x; jj < i; jj += kWave) q += hsh[jj] * hsh[jj];
        for (int o = kWave >> 1; o > 0; o >>= 1) q += __shfl_xor(q, o, kWave);
        if (threadIdx.x == 0) inv_s = q;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double hii = sqrt(fabs(hsh[i] - inv_s));
        hsh[i] = hii;
        inv_s = 1.0 / hii;
    }
    __syncthreads();
    const double inv = inv_s;
    // one element per lane (coalesced 4- or 8-byte loads across the wave):
    // four times the blocks of a vector-load layout, so the latency-bound
    // sweep has a wave on every SIMD; the basis loads go out 16 at a time
    BT* w = basis + (size_t)i * n;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        double v = (double)w[q];
        int jj = 0;
        for (; jj + 16 <= i; jj += 16) {
            BT u[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) u[m] = basis[(size_t)(jj + m) * n + q];
#pragma unroll
            for (int m = 0; m < 16; ++m) v = v + hc[jj + m] * (double)u[m];
        }
        for (; jj + 4 <= i; jj += 4) {
            BT u[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) u[m] = basis[(size_t)(jj + m) * n + q];
#pragma unroll
            for (int m = 0; m < 4; ++m) v = v + hc[jj + m] * (double)u[m];
        }
        for (; jj < i; ++jj) v = v + hc[jj] * (double)basis[(size_t)jj * n + q];
        w[q] = (BT)(v * inv);
    }
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    st->iters += 1;
    {
        double* Hc = A.Hu + (size_t)(i - 1) * (sD + 1); // unrotated column i
        for (int jj = 0; jj <= i; ++jj) Hc[jj] = hsh[jj];
    }
    for (int jj = 1; jj <= i - 1; ++jj) {
        const double c = cs_s[0][jj - 1], sn = cs_s[1][jj - 1];
        const double tmp = c * hsh[jj - 1] + sn * hsh[jj];
        hsh[jj] = -sn * hsh[jj - 1] + c * hsh[jj];
        hsh[jj - 1] = tmp;
    }
    const double tmp = sqrt(hsh[i - 1] * hsh[i - 1] + hsh[i] * hsh[i]);
    const double ci = hsh[i - 1] / tmp;
    const double si = hsh[i] / tmp;
    hsh[i - 1] = tmp;
    hsh[i] = 0.0;
    A.c[i - 1] = ci;
    A.sn[i - 1] = si;
    double* Hc = A.H + (size_t)(i - 1) * (sD + 1);
    for (int jj = 0; jj <= i; ++jj) Hc[jj] = hsh[jj];
    const double ei = A.err[i - 1];
    const double eNew = -si * ei;
    A.err[i] = eNew;
    A.err[i - 1] = ci * ei;
    if (fabs(eNew) < st->eps) {
        st->suc = 1;
        st->stop = 1;
        st->iconv = i;
    } else if (i >= sD) {
        st->stop = 1;
        st->iconv = sD;
    }
}

// Inner cycle, dof 4, FP32: the SpMV w = K u_{i-1} (written to basis slot
// i) and, in the same pass, per-block partials of <u_j, w> for j < i and of
// <w, w> (slot i), so the step needs no separate multi-dot sweep.  Lane
// layout as k_bspmv4<float, 4>: 16 lanes per block row, lane (h, c) with
// c the output component; after the row sum every lane holds w(row, c), and
// lane (h, c) forms the products of the coefficients j = h, h + 4, ...
// Supports i + 1 <= 32 coefficients.
template <bool NT>
__global__ __launch_bounds__(256) void k_gm_spmv_dots4(int nNo, const int* __restrict__ rowPtr,
                                                       const int* __restrict__ colPtr,
                                                       const float* __restrict__ val, float* __restrict__ basis,
                                                       size_t n, int i, double* __restrict__ partial, int nb,
                                                       const GmState* __restrict__ st)
{
    if (st->stop) return;
    __shared__ double red[4][32];
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int row = t >> 4;
    const int l = t & 15;
    const int h = l >> 2, c = l & 3;
    const float* x = basis + (size_t)(i - 1) * n;
    double a0 = 0.0, a1 = 0.0;
    if (row < nNo) {
        const int e = rowPtr[row + 1];
        int k = rowPtr[row] + h;
        for (; k + 4 < e; k += 8) {
            const int c0 = colPtr[k], c1 = colPtr[k + 4];
            a0 += dot4_nt<NT>(val + (size_t)k * 16 + c * 4, x + (size_t)c0 * 4);
            a1 += dot4_nt<NT>(val + (size_t)(k + 4) * 16 + c * 4, x + (size_t)c1 * 4);
        }
        if (k < e) a0 += dot4_nt<NT>(val + (size_t)k * 16 + c * 4, x + (size_t)colPtr[k] * 4);
    }
    double acc = a0 + a1;
    acc += __shfl_xor(acc, 4, kWave);
    acc += __shfl_xor(acc, 8, kWave);
    const float wv = (float)acc;
    const size_t q = (size_t)row * 4 + c;
    if (row < nNo && h == 0) basis[(size_t)i * n + q] = wv;
    const double wd = (double)wv;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int j = h + 4 * m;
        double pv = 0.0;
        if (row < nNo && j <= i) pv = (j == i ? wd : (double)basis[(size_t)j * n + q]) * wd;
        // sum over the four components and the four rows of the wavefront
        pv += __shfl_xor(pv, 1, kWave);
        pv += __shfl_xor(pv, 2, kWave);
        pv += __shfl_xor(pv, 16, kWave);
        pv += __shfl_xor(pv, 32, kWave);
        if ((lane & 0x33) == 0 && j <= i) red[wave][j] = pv;
    }
    __syncthreads();
    if (threadIdx.x <= i)
        partial[(size_t)threadIdx.x * nb + blockIdx.x] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// h(j,i), j <= i: the per-block partials of k_gm_spmv_dots4 summed in block
// order, one block per coefficient, into A.hraw.
__global__ __launch_bounds__(256) void k_gm_hsum(int i, int nb, const double* __restrict__ partial,
                                                 const GmState* __restrict__ st, GmArrays A)
{
    if (st->stop) return;
    __shared__ double lds[4];
    const double* pp = partial + (size_t)blockIdx.x * nb;
    double t = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) t += pp[b];
    t = block_sum<256>(t, lds);
    if (threadIdx.x == 0) A.hraw[blockIdx.x] = t;
}

// Back substitution of the cycle (GMRES.f:369-376).  The block stages the
// upper triangle of H in LDS (packed by columns) when it fits, then one
// thread runs the reference's row-oriented substitution.
constexpr int kBackLds = 6144; // doubles of LDS for the packed triangle (48 KiB, sD <= 110)

__global__ __launch_bounds__(64) void k_gm_backsolve(int sD, GmState* __restrict__ st, GmArrays A)
{
    // One wavefront.  Lane l owns rows l, l+64, ...; columns are eliminated
    // from the last one: y_j = acc_j / h(j,j), then every owner of a row
    // r < j subtracts h(r,j) y_j (the column-oriented form of GMRES.f:370-376;
    // each row accumulates the same products, in descending column order).
    extern __shared__ double Hs[];
    int i = st->iconv;
    if (i > sD) i = sD;
    const bool lds = (size_t)sD * (sD + 1) / 2 <= (size_t)kBackLds;
    const int lane = threadIdx.x;
    if (lds) {
        for (int k = 0; k < i; ++k)
            for (int r = lane; r <= k; r += 64) Hs[k * (k + 1) / 2 + r] = A.H[(size_t)k * (sD + 1) + r];
    }
    __syncthreads();
    auto H = [&](int r, int k) -> double {  // 0-based h(r+1,k+1), r <= k
        return lds ? Hs[k * (k + 1) / 2 + r] : A.H[(size_t)k * (sD + 1) + r];
    };
    constexpr int kRows = (kMaxSD + 63) / 64;
    double acc[kRows];
#pragma unroll
    for (int m = 0; m < kRows; ++m) {
        const int r = lane + 64 * m;
        acc[m] = (r < i) ? A.err[r] : 0.0;
    }
    for (int j = i - 1; j >= 0; --j) {
        const int owner = j & 63, mj = j >> 6;
        double yj = 0.0;
#pragma unroll
        for (int m = 0; m < kRows; ++m)
            if (m == mj) yj = acc[m] / H(j, j);
        yj = __shfl(yj, owner, kWave);
        if (lane == owner) {
#pragma unroll
            for (int m = 0; m < kRows; ++m)
                if (m == mj) acc[m] = yj;
        }
#pragma unroll
        for (int m = 0; m < kRows; ++m) {
            const int r = lane + 64 * m;
            if (r < j) acc[m] = acc[m] - H(r, j) * yj;
        }
    }
#pragma unroll
    for (int m = 0; m < kRows; ++m) {
        const int r = lane + 64 * m;
        if (r < i) A.y[r] = acc[m];
    }
    if (lane == 0) st->fNorm = fabs(A.err[i]);
}

// X = X + y_j u_j, j = 1..i (GMRES.f:378-381)
template <typename BT>
__global__ __launch_bounds__(256) void k_gm_xupdate(size_t n, BT* __restrict__ X, const BT* __restrict__ basis,
                                                    const GmState* __restrict__ st, GmArrays A)
{
    const int i = st->iconv;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        double v = (double)X[q];
        for (int jj = 0; jj < i; ++jj) v = v + A.y[jj] * (double)basis[(size_t)jj * n + q];
        X[q] = (BT)v;
    }
}

template <typename BT>
__global__ void k_set_zero(size_t n, BT* __restrict__ x)
{
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) x[q] = (BT)0;
}

template <typename BT>
__global__ __launch_bounds__(256) void k_normsq_partial(size_t n, const BT* __restrict__ u,
                                                        double* __restrict__ partial)
{
    __shared__ double lds[4];
    double s = 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        s += (double)u[q] * (double)u[q];
    s = block_sum<256>(s, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_sum_to(int m, const double* __restrict__ partial, double* __restrict__ out)
{
    __shared__ double lds[4];
    double s = 0.0;
    for (int q = threadIdx.x; q < m; q += 256) s += partial[q];
    s = block_sum<256>(s, lds);
    if (threadIdx.x == 0) *out = s;
}

// Generic operator: y = A x on the device (FP64 or FP32 matrix, plain or
// transposed), optionally gated by the device stop flag.
template <typename BT> struct GmOp;

template <> struct GmOp<double> {
    Lhs* L; int dof; const double* K; bool transpose;
    int cpl = 0;  // coupled faces: 0 none, 1 ADDBCMUL ADD, 2 ADD then PRE (GMRES.f:109-115)
    int apply(const double* x, double* y, hipStream_t s, const int* stop) const {
        SV_TRY(transpose ? launch_spmv_t(*L, dof, K, x, y, s, stop) : launch_spmv(*L, dof, K, x, y, s, stop));
        if (cpl >= 1) SV_TRY(cpl_apply(*L, 0, dof, x, y, s, stop));
        if (cpl >= 2) SV_TRY(cpl_apply(*L, 1, dof, y, y, s, stop));
        return SVFSI_OK;
    }
};
template <> struct GmOp<float> {
    const Lhs* L; int dof; const float* K;
    const float* dinv; // optional block-Jacobi: the operator is D^-1 K
    int apply(const float* x, float* y, hipStream_t s, const int* stop) const {
        return launch_spmv_f32(*L, dof, K, x, y, s, stop, dinv);
    }
};

static bool op_transposed(const GmOp<double>& op) { return op.transpose; }
static bool op_transposed(const GmOp<float>&) { return false; }
static int op_cpl(const GmOp<double>& op) { return op.cpl; }
static int op_cpl(const GmOp<float>&) { return 0; }


// Restarted GMRES(sD) on n = nNo*dof unknowns.  R: RHS in, solution out
// (GMRESV overwrites R with X, GMRES.f:383).  Fills suc/itr/iNorm/fNorm/dB.
template <typename BT>
static int gmres_core(Lhs& L, const GmOp<BT>& op, int dof, BT* R, DBuf& basisBuf, DBuf& xBuf,
                      FSILS_subLsType& ls, hipStream_t s)
{
    const size_t n = (size_t)L.nNo * dof;
    const int sD = ls.sD > 0 ? ls.sD : 1;
    if (sD > kMaxSD) {
        set_error("GMRES: Krylov dimension %d exceeds %d", sD, kMaxSD);
        return SVFSI_ERR_NOTSUP;
    }
    Work& w = L.ws;
    // row chunks of the multi-dot: 256 lanes x E elements each
    const size_t nbr = (n + (size_t)256 * VecE<BT>::E - 1) / ((size_t)256 * VecE<BT>::E);
    const long long foldAt = env_int("SVFSI_FOLD_CHUNKS", kFoldChunks);
    const unsigned gv = vec_grid(n);
    SV_TRY(basisBuf.ensure((size_t)(sD + 1) * n * sizeof(BT)));
    SV_TRY(xBuf.ensure(n * sizeof(BT)));
    const size_t partN = ((size_t)(sD + 2) * nbr > gv ? (size_t)(sD + 2) * nbr : gv) + 16;
    SV_TRY(w.partial.ensure(partN * sizeof(double)));
    SV_TRY(w.state.ensure(256));
    SV_TRY(w.hess.ensure(gm_arrays_doubles(sD) * sizeof(double)));
    BT* basis = basisBuf.as<BT>();
    BT* X = xBuf.as<BT>();
    double* part = w.partial.as<double>();
    double* scal = part + partN - 8;
    GmState* st = w.state.as<GmState>();
    GmArrays A = gm_carve(w.hess.as<double>(), sD);
    const int* stopFlag = &st->stop;

    ls.callD = wall_seconds();
    ls.suc = 0;
    hipLaunchKernelGGL(k_normsq_partial<BT>, dim3(gv), dim3(256), 0, s, n, (const BT*)R, part);
    hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, s, (int)gv, part, scal);
    double rr = 0.0;
    SV_HIP(hipMemcpyAsync(&rr, scal, sizeof(double), hipMemcpyDeviceToHost, s));
    SV_HIP(hipStreamSynchronize(s));
    double eps = std::sqrt(rr);
    ls.iNorm = eps;
    ls.fNorm = eps;
    eps = std::fmax(ls.absTol, ls.relTol * eps);
    ls.itr = 0;
    ls.cM = 0;
    hipLaunchKernelGGL(k_set_zero<BT>, dim3(gv), dim3(256), 0, s, n, X);
    if (ls.iNorm <= ls.absTol) {
        // GMRES.f:307-311: return before R = X (R keeps the scaled RHS)
        ls.callD = 2.220446049250313e-16;
        ls.dB = 0.0;
        return SVFSI_OK;
    }
    hipLaunchKernelGGL(k_gm_init, dim3(1), dim3(64), 0, s, st, eps);

    PinnedPoll& poll = w.poll;
    SV_TRY(poll.ensure(256));
    GmState hs{};
    const int chunk = 8;
    const bool useGraphs = graphs_enabled();
    if (useGraphs) {
        std::vector<uintptr_t> sig = {
            (uintptr_t)n, (uintptr_t)dof, (uintptr_t)sD, (uintptr_t)op.K, (uintptr_t)sizeof(BT),
            (uintptr_t)op_transposed(op), (uintptr_t)basis, (uintptr_t)part, (uintptr_t)st, (uintptr_t)w.hess.p,
            (uintptr_t)L.rowPtr.p, (uintptr_t)L.colPtr.p, (uintptr_t)L.trans.p, (uintptr_t)op_cpl(op), (uintptr_t)((long long)nbr > foldAt)};
        if (op_cpl(op)) cpl_sig(L, sig);
        if (sig != w.gmGraphs.sig) {
            w.gmGraphs.clear();
            w.gmGraphs.sig = sig;
        }
    }
    for (int l = 1; l <= ls.mItr; ++l) {
        ls.dB = ls.fNorm;
        ls.itr += 1;
        // u1 = R - K X ; err(1) = |u1| ; u1 = u1/err(1)
        SV_TRY(op.apply(X, basis, s, nullptr));
        hipLaunchKernelGGL(k_gm_resid<BT>, dim3(gv), dim3(256), 0, s, n, (const BT*)R, basis, part, st, A);
        hipLaunchKernelGGL(k_gm_div<BT>, dim3(gv), dim3(256), 0, s, n, basis, A);
        ls.cM += 1;
        int posted = 0;
        bool stopped = false;
        for (int i0 = 1; i0 <= sD && !stopped; i0 += chunk) {
            const int i1 = (i0 + chunk - 1 < sD) ? i0 + chunk - 1 : sD;
            // Arnoldi steps i0..i1: the same launches every restart cycle,
            // replayed as one graph per chunk
            auto steps = [&, i0, i1](hipStream_t cs) -> int {
                for (int i = i0; i <= i1; ++i) {
                    SV_TRY(op.apply(basis + (size_t)(i - 1) * n, basis + (size_t)i * n, cs, stopFlag));
                    hipLaunchKernelGGL(k_gm_mdot<BT>, dim3((unsigned)nbr, (unsigned)((i + 8) / 8)), dim3(256), 0, cs,
                                       n, i, (const BT*)basis, part, (const GmState*)st);
                    if ((long long)nbr > foldAt) {
                        hipLaunchKernelGGL(k_gm_fold, dim3((unsigned)(i + 1)), dim3(256), 0, cs, (int)nbr, part,
                                           (const GmState*)st);
                        hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(256), 0, cs, i, sD, 1, (int)nbr,
                                           (const double*)part, st, A, 0);
                    } else {
                        hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(256), 0, cs, i, sD, (int)nbr, (int)nbr,
                                           (const double*)part, st, A, 0);
                    }
                    hipLaunchKernelGGL(k_gm_update<BT>, dim3(gv), dim3(256), 0, cs, n, i, sD, basis,
                                       (const GmState*)st, A, nullptr);
                }
                SV_HIP(hipGetLastError());
                return SVFSI_OK;
            };
            if (useGraphs)
                SV_TRY(graph_run(w.gmGraphs, i0, s, steps));
            else
                SV_TRY(steps(s));
            const int q = posted & 1;
            SV_TRY(poll.post(q, st, sizeof(GmState), s));
            ++posted;
            if (posted >= 2) {
                SV_TRY(poll.wait(q ^ 1));
                std::memcpy(&hs, poll.h[q ^ 1], sizeof(GmState));
                if (hs.stop) stopped = true;
            }
        }
        {
            const size_t tri = (size_t)sD * (sD + 1) / 2;
            const size_t ldsB = (tri <= (size_t)kBackLds ? tri : 0) * sizeof(double);
            hipLaunchKernelGGL(k_gm_backsolve, dim3(1), dim3(64), ldsB, s, sD, st, A);
        }
        hipLaunchKernelGGL(k_gm_xupdate<BT>, dim3(gv), dim3(256), 0, s, n, X, (const BT*)basis, (const GmState*)st, A);
        SV_HIP(hipGetLastError());
        SV_HIP(hipMemcpyAsync(&hs, st, sizeof(GmState), hipMemcpyDeviceToHost, s));
        SV_HIP(hipStreamSynchronize(s));
        ls.itr += hs.iconv;
        ls.cM += hs.iconv;
        ls.fNorm = hs.fNorm;
        if (hs.suc) {
            ls.suc = 1;
            break;
        }
    }
    SV_HIP(hipMemcpyAsync(R, X, n * sizeof(BT), hipMemcpyDeviceToDevice, s));
    SV_HIP(hipStreamSynchronize(s));
    ls.callD = wall_seconds() - ls.callD;
    ls.dB = 10.0 * std::log(ls.fNorm / ls.dB);
    return SVFSI_OK;
}

int gmres_solve(Lhs& L, FSILS_subLsType& ls, int dof, const double* K, double* R, hipStream_t s,
                bool transpose, int cpl)
{
    GmOp<double> op{&L, dof, K, transpose, cpl};
    return gmres_core<double>(L, op, dof, R, L.ws.basis, L.ws.X, ls, s);
}

// ---------------------------------------------------------------------------
// Device-driven GMRES cycle with zero initial guess (the inner solver of the
// flexible GMRES): no host round trip.  The relative tolerance is applied on
// the device to |R|; `gate` (an outer stop flag) cancels the whole cycle.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_gm_start(int m, const double* __restrict__ partial,
                                                  GmState* __restrict__ st, GmArrays A, double relTol,
                                                  const int* __restrict__ gate)
{
    __shared__ double lds[4];
    double t = 0.0;
    for (int q = threadIdx.x; q < m; q += 256) t += partial[q];
    t = block_sum<256>(t, lds);
    if (threadIdx.x == 0) {
        const double e0 = sqrt(t);
        A.err[0] = e0;
        st->eps = relTol * e0;
        st->iconv = 0;
        st->suc = 0;
        st->stop = ((gate && *gate) || e0 == 0.0) ? 1 : 0;
    }
}

template <typename BT>
__global__ __launch_bounds__(256) void k_gm_scale_copy(size_t n, const BT* __restrict__ R, BT* __restrict__ u0,
                                                       const GmState* __restrict__ st, GmArrays A)
{
    if (st->stop) return;
    const double e0 = A.err[0];
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        u0[q] = (BT)((double)R[q] / e0);
}

// R = sum_{j < iconv} y_j u_j (zero initial guess)
template <typename BT>
__global__ __launch_bounds__(256) void k_gm_xsolve(size_t n, BT* __restrict__ R, const BT* __restrict__ basis,
                                                   const GmState* __restrict__ st, GmArrays A)
{
    const int i = st->iconv;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        double v = 0.0;
        for (int jj = 0; jj < i; ++jj) v = v + A.y[jj] * (double)basis[(size_t)jj * n + q];
        R[q] = (BT)v;
    }
}

// Partial-sum slots of the inner cycle: the multi-dot (row chunks of 256
// lanes x 4 elements) or the fused SpMV + dot (16 rows per block), and the
// vector grid of the norm.
static size_t inner_partials(int nNo, size_t n, int sD, size_t nbr, unsigned gv)
{
    size_t m = (size_t)(sD + 2) * nbr;
    const size_t f = (size_t)(sD + 2) * (size_t)((nNo + 15) / 16);
    if (f > m) m = f;
    if (gv > m) m = gv;
    return m + 16;
}

// The fused SpMV + multi-dot inner step is off unless SVFSI_FUSED_INNER=1:
// measured on the 200k-tet cylinder it is slower (14.1 us + 4.1 us for the
// coefficient sums) than the SpMV and multi-dot it replaces (8.9 + 6.2 us).
static bool fused_inner_enabled()
{
    const char* v = getenv("SVFSI_FUSED_INNER");
    return v && v[0] == '1';
}

// One cycle of GMRES(sD) on R (RHS in, approximate solution out), enqueued
// without any host synchronisation.  Uses the workspace's inner state.
template <typename BT>
static int gmres_inner_dev(Lhs& L, const GmOp<BT>& op, int dof, BT* R, DBuf& basisBuf, int sD, double relTol,
                           const int* gate, hipStream_t s)
{
    const size_t n = (size_t)L.nNo * dof;
    Work& w = L.ws;
    const size_t nbr = (n + (size_t)256 * VecE<BT>::E - 1) / ((size_t)256 * VecE<BT>::E);
    const unsigned gv = vec_grid(n);
    SV_TRY(basisBuf.ensure((size_t)(sD + 1) * n * sizeof(BT)));
    const size_t partN = inner_partials(L.nNo, n, sD, nbr, gv);
    SV_TRY(w.partial.ensure(partN * sizeof(double)));
    SV_TRY(w.hess.ensure(gm_arrays_doubles(sD) * sizeof(double)));
    BT* basis = basisBuf.as<BT>();
    double* part = w.partial.as<double>();
    GmState* st = w.state.as<GmState>();
    GmArrays A = gm_carve(w.hess.as<double>(), sD);
    hipLaunchKernelGGL(k_normsq_partial<BT>, dim3(gv), dim3(256), 0, s, n, (const BT*)R, part);
    hipLaunchKernelGGL(k_gm_start, dim3(1), dim3(256), 0, s, (int)gv, (const double*)part, st, A, relTol, gate);
    hipLaunchKernelGGL(k_gm_scale_copy<BT>, dim3(gv), dim3(256), 0, s, n, (const BT*)R, basis, (const GmState*)st, A);
    bool fused = false;
    if constexpr (std::is_same<BT, float>::value) {
        if (dof == 4 && sD + 1 <= 32 && fused_inner_enabled() && !op.dinv) {
            fused = true;
            // fused SpMV + multi-dot, coefficient sums, update: three
            // launches per Arnoldi step
            const int nbS = (L.nNo + 15) / 16;
            const bool nt = (size_t)L.nnz * 16 * sizeof(float) > ((size_t)128 << 20);
            const double spmvBytes = (double)L.nnz * 64.0 + 4.0 * L.nnz + 4.0 * (L.nNo + 1) + 2.0 * n * 4.0;
            for (int i = 1; i <= sD; ++i) {
                // SpMV bytes + the i basis vectors the dots read + partials
                ProfScope ps(1, spmvBytes + (double)i * n * 4.0 + (double)(i + 1) * nbS * 8.0);
                if (nt)
                    SV_LAUNCH(ps, k_gm_spmv_dots4<true>, dim3((unsigned)nbS), dim3(256), 0, s, L.nNo,
                              L.rowPtr.as<int>(), L.colPtr.as<int>(), op.K, basis, n, i, part, nbS,
                              (const GmState*)st);
                else
                    SV_LAUNCH(ps, k_gm_spmv_dots4<false>, dim3((unsigned)nbS), dim3(256), 0, s, L.nNo,
                              L.rowPtr.as<int>(), L.colPtr.as<int>(), op.K, basis, n, i, part, nbS,
                              (const GmState*)st);
                hipLaunchKernelGGL(k_gm_hsum, dim3((unsigned)(i + 1)), dim3(256), 0, s, i, nbS, (const double*)part,
                                   (const GmState*)st, A);
                hipLaunchKernelGGL((k_gm_rotupdate<BT, true>), dim3(gv), dim3(256), 0, s, n, i, sD, nbS,
                                   (const double*)part, basis, st, A);
            }
        }
    }
    for (int i = 1; i <= sD && !fused; ++i) {
        SV_TRY(op.apply(basis + (size_t)(i - 1) * n, basis + (size_t)i * n, s, &st->stop));
        hipLaunchKernelGGL(k_gm_mdot<BT>, dim3((unsigned)nbr, (unsigned)((i + 8) / 8)), dim3(256), 0, s, n, i,
                           (const BT*)basis, part, (const GmState*)st);
        hipLaunchKernelGGL(k_gm_rotupdate<BT>, dim3(gv), dim3(256), 0, s, n, i, sD, (int)nbr,
                           (const double*)part, basis, st, A);
    }
    {
        const size_t tri = (size_t)sD * (sD + 1) / 2;
        const size_t ldsB = (tri <= (size_t)kBackLds ? tri : 0) * sizeof(double);
        hipLaunchKernelGGL(k_gm_backsolve, dim3(1), dim3(64), ldsB, s, sD, st, A);
    }
    hipLaunchKernelGGL(k_gm_xsolve<BT>, dim3(gv), dim3(256), 0, s, n, R, (const BT*)basis, (const GmState*)st, A);
    SV_HIP(hipGetLastError());
    return SVFSI_OK;
}

// b = a converted between the outer (FP64) and inner (IT) precisions
template <typename TI, typename TO>
__global__ void k_cvt(size_t n, const TI* __restrict__ a, TO* __restrict__ b)
{
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) b[q] = (TO)a[q];
}

// Start of an outer step's preconditioner application, one pass per node:
// p = D^-1 (IT) q (p = (IT) q without dinv; the product summed in FP64 as
// k_blockdiag_apply does) and, when z is given, z = 0.
template <typename IT, int DOF>
__global__ __launch_bounds__(256) void k_fg_start(int nNo, const double* __restrict__ q, const IT* __restrict__ dinv,
                                                  IT* __restrict__ p, IT* __restrict__ z)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nNo) return;
    IT xi[DOF];
#pragma unroll
    for (int m = 0; m < DOF; ++m) xi[m] = (IT)q[(size_t)r * DOF + m];
    if (dinv) {
        const IT* d = dinv + (size_t)r * DOF * DOF;
#pragma unroll
        for (int a = 0; a < DOF; ++a) {
            double v = 0.0;
#pragma unroll
            for (int m = 0; m < DOF; ++m) v += (double)d[a * DOF + m] * (double)xi[m];
            p[(size_t)r * DOF + a] = (IT)v;
        }
    } else {
#pragma unroll
        for (int m = 0; m < DOF; ++m) p[(size_t)r * DOF + m] = xi[m];
    }
    if (z)
#pragma unroll
        for (int m = 0; m < DOF; ++m) z[(size_t)r * DOF + m] = (IT)0;
}

template <typename IT>
static void launch_fg_start(int nNo, int dof, const double* q, const IT* dinv, IT* p, IT* z, hipStream_t s)
{
    const unsigned g = (unsigned)((nNo + 255) / 256);
    if (dof == 4) hipLaunchKernelGGL((k_fg_start<IT, 4>), dim3(g), dim3(256), 0, s, nNo, q, dinv, p, z);
    else if (dof == 3) hipLaunchKernelGGL((k_fg_start<IT, 3>), dim3(g), dim3(256), 0, s, nNo, q, dinv, p, z);
    else if (dof == 2) hipLaunchKernelGGL((k_fg_start<IT, 2>), dim3(g), dim3(256), 0, s, nNo, q, dinv, p, z);
    else hipLaunchKernelGGL((k_fg_start<IT, 1>), dim3(g), dim3(256), 0, s, nNo, q, dinv, p, z);
}

// The last, linear factor of the polynomial (y += c1 prod, as k_poly_last)
// with the FP64 copy of the result the outer product reads.
template <typename IT>
__global__ __launch_bounds__(256) void k_poly_last_cvt(size_t n, const IT* __restrict__ prod, IT* __restrict__ y,
                                                       double* __restrict__ y64, const double* __restrict__ coef,
                                                       int step)
{
    const double c1 = coef[2 * step];
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        const IT v = (IT)((double)y[q] + c1 * (double)prod[q]);
        y[q] = v;
        y64[q] = (double)v;
    }
}

// x += sum_j y_j z_j with the inner-precision directions z_j (flexible
// GMRES update)
template <typename IT>
__global__ __launch_bounds__(256) void k_fg_xupdate(size_t n, double* __restrict__ X, const IT* __restrict__ Z,
                                                    const GmState* __restrict__ st, GmArrays A)
{
    const int i = st->iconv;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        double v = X[q];
        for (int jj = 0; jj < i; ++jj) v = v + A.y[jj] * (double)Z[(size_t)jj * n + q];
        X[q] = v;
    }
}

static double env_dbl(const char* k, double d)
{
    const char* e = getenv(k);
    return e ? atof(e) : d;
}

// Mixed precision: flexible GMRES (FGMRES(m)) in FP64 whose right
// preconditioner is an inner GMRES on the FP32 copy of the scaled matrix.
// FP64: the outer residual b - A x, the outer Arnoldi (classical
// Gram-Schmidt with the Pythagorean norm, as GMRESV), the Hessenberg /
// Givens least squares and the solution x.  FP32: the inner matrix, SpMV,
// basis storage and the flexible directions z_j (the inner solver's dot
// products and Hessenberg arithmetic are FP64).  The inner solve stops at
// a loose relative tolerance; the outer Krylov space keeps what restarted
// inner cycles would lose.  Convergence is GMRESV's FP64 test
// |W(R - K x)| < max(absTol, relTol |W R|) on the outer estimate.
//
// IT is the inner precision: float for the mixed solve, double for the same
// nested algorithm run entirely in FP64 (the comparison that separates what
// the algorithm gains from what FP32 storage gains).  Kin is the inner
// matrix: the FP32 copy of the scaled matrix written by the diagonal
// preconditioner (IT = float), or nullptr (IT = double: the scaled FP64
// matrix itself, or a copy in w.valIn when block Jacobi rescales it).
template <typename IT>
static int nested_solve(Lhs& L, FSILS_subLsType& ls, int dof, const double* K, IT* Kin, double* R, hipStream_t s)
{
    const size_t n = (size_t)L.nNo * dof;
    Work& w = L.ws;
    // outer Krylov dimension: the caller's restart length (dimKry), 30 when
    // none is set (CG requests); SVFSI_FGMRES_M overrides
    const int mo = env_int("SVFSI_FGMRES_M", ls.sD > 0 ? std::min(ls.sD, kMaxSD) : 30);
    const int mi = env_int("SVFSI_FGMRES_INNER_M", 30);     // inner Krylov dimension
    if (mo > kMaxSD || mi > kMaxSD) {
        set_error("FGMRES: Krylov dimension exceeds %d", kMaxSD);
        return SVFSI_ERR_NOTSUP;
    }
    const unsigned gv = vec_grid(n);
    const size_t nbr = (n + (size_t)256 * VecE<double>::E - 1) / ((size_t)256 * VecE<double>::E);
    const long long foldAt = env_int("SVFSI_FOLD_CHUNKS", kFoldChunks);
    SV_TRY(w.X.ensure(n * sizeof(double)));
    SV_TRY(w.tmp.ensure(n * sizeof(double)));
    SV_TRY(w.basis.ensure((size_t)(mo + 1) * n * sizeof(double)));
    SV_TRY(w.zdir.ensure((size_t)mo * n * sizeof(IT)));
    const size_t partN = ((size_t)(mo + 2) * nbr > gv ? (size_t)(mo + 2) * nbr : gv) + 16;
    SV_TRY(w.opartial.ensure(partN * sizeof(double)));
    SV_TRY(w.ostate.ensure(256));
    SV_TRY(w.ohess.ensure(gm_arrays_doubles(mo) * sizeof(double)));
    double* x = w.X.as<double>();
    double* t64 = w.tmp.as<double>();
    double* Q = w.basis.as<double>();
    IT* Z = w.zdir.as<IT>();
    double* part = w.opartial.as<double>();
    double* scal = part + partN - 8;
    GmState* st = w.ostate.as<GmState>();
    GmArrays A = gm_carve(w.ohess.as<double>(), mo);
    // Inner work: the polynomial preconditioner (dof 3, 4; default) or one
    // FP32 GMRES cycle (SVFSI_INNER_POLY=0).  Block-Jacobi (inner operator
    // D^-1 K, D the dof x dof diagonal blocks of the scaled matrix) is on by
    // default for dof 4: on the fluid cylinders it cut the inner GMRES
    // cycle's SpMVs (467 -> 405) and the polynomial's outer steps on the
    // 3.2M-tet cylinder (72 -> 54).  Off for dof 3, where it added outer
    // steps on the elasticity cube (16 -> 22).  SVFSI_INNER_BJ = 0 / 1
    // overrides.
    // A saddle-point system without pressure stabilisation (unmasked rows
    // with a zero diagonal: the Taylor-Hood fluid) defeats both the node
    // block inverses and the fixed polynomial: on the tet10 channel they
    // made the outer loop stall or overflow, while one inner GMRES cycle of
    // the diagonally scaled operator converged (371 outer steps at 10^3).
    // Such systems get the inner GMRES cycle without block Jacobi unless
    // SVFSI_INNER_POLY / SVFSI_INNER_BJ say otherwise.
    int zeroDiag = 0;
    SV_TRY(w.ostate.ensure(256));
    SV_TRY(count_zero_diag(L, dof, K, w.W.as<double>(), w.ostate.as<int>(), zeroDiag, s));
    const bool saddle = zeroDiag > 0;
    // outer Gram-Schmidt (SVFSI_OUTER_CGS): 3 a single classical pass with
    // the explicit norm (the default), 2 two passes (CGS2, the default on
    // saddle-point systems), 1 the reference's single pass with the
    // Pythagorean norm.  Measured (outer steps, ms per mixed step):
    // 200k-tet cylinder 22 / 6.64 (3) against 22 / 6.77 (2) and 54 / 17
    // (1, before the later kernels); 3.2M-tet cylinder 89 / 139.5 against
    // 89 / 161.2; LELAS 20^3 18 steps either way; the Taylor-Hood channel
    // 390 / 504 against 354 / 460, where the second pass pays.  The
    // Pythagorean h(i+1,i) loses the basis' orthogonality on these
    // operators; the explicit norm keeps it as well as the second pass does
    // on the others, at half the basis reads.
    const int cgsMode = [saddle] {
        const int v = env_int("SVFSI_OUTER_CGS", saddle ? 2 : 3);
        return (v == 1 || v == 2) ? v : 3;
    }();
    const int cgsPasses = cgsMode == 2 ? 2 : 1;
    // the explicit-norm pass's end as one kernel (k_gm_normscale; the same
    // arithmetic as k_gm_givens pass 3 then k_gm_scale, SVFSI_GM_NORMSCALE=0)
    const bool fuseNorm = env_int("SVFSI_GM_NORMSCALE", 1) != 0;
    // One V-cycle of the aggregation multigrid (amg.hip) on the FP32 scaled
    // matrix (no block-Jacobi rescaling: the smoother uses the node blocks
    // itself): the default for dof-4 systems whose FP32 matrix exceeds half
    // the Infinity Cache, where the polynomial's outer steps grow with the
    // mesh (3.2M-tet cylinder: 177 ms, 89 outer steps and 446 fine SpMVs
    // against 227-233 ms at degree 30 and 216 ms at 50; 200k-tet cylinder
    // 15-16 ms against 6.7-7.0 for the polynomial).  SVFSI_INNER_AMG = 0 / 1
    // overrides.
    const bool bigFluid = dof == 4 && std::is_same<IT, float>::value && !saddle &&
                          (size_t)L.nnz * 16 * sizeof(float) > ((size_t)128 << 20);
    const bool useAmg = (dof == 3 || dof == 4) && env_int("SVFSI_INNER_AMG", bigFluid ? 1 : 0) != 0;
    if (useAmg && !std::is_same<IT, float>::value) {
        set_error("the multigrid inner preconditioner is FP32 only");
        return SVFSI_ERR_NOTSUP;
    }
    // inner GMRES tolerance: 0.1; 0.01 on saddle-point systems (the P2P1
    // channel at 10^3: 227 outer steps and 138 ms against 371 and 197 ms)
    const double tin = env_dbl("SVFSI_FGMRES_INNER_TOL", saddle ? 0.01 : 0.1);
    const bool polyWanted =
        !useAmg && (dof == 3 || dof == 4) && env_int("SVFSI_INNER_POLY", saddle ? 0 : 1) != 0;
    const bool bj =
        !useAmg && (dof == 3 || dof == 4) && env_int("SVFSI_INNER_BJ", (dof == 4 && !saddle) ? 1 : 0) != 0;
    // With block Jacobi the inner matrix is replaced by B = D^-1 K (formed
    // once per solve from the FP64 matrix into Kin, or into w.valIn for an
    // FP64 inner operator), so the inner products carry no epilogue, and
    // each inner right-hand side is multiplied by D^-1.
    const IT* dinv = nullptr;
    if (!Kin) {
        if constexpr (std::is_same<IT, double>::value) {
            if (bj) {
                SV_TRY(w.valIn.ensure((size_t)L.nnz * dof * dof * sizeof(double)));
                Kin = w.valIn.as<double>();
            } else {
                Kin = const_cast<double*>(K); // read only: no block Jacobi
            }
        } else {
            set_error("nested solve: FP32 inner matrix missing");
            return SVFSI_ERR_ARG;
        }
    }
    if (bj) {
        SV_TRY(w.dinv32.ensure((size_t)L.nNo * dof * dof * sizeof(IT)));
        SV_TRY(launch_blockdiag_inv<IT>(L, dof, K, w.dinv32.as<IT>(), s));
        SV_TRY(launch_bjscale<IT>(L, dof, K, w.dinv32.as<IT>(), Kin, s));
        dinv = w.dinv32.as<IT>();
    }
    GmOp<IT> opIn{}; // no transpose, no coupled faces, no D^-1 epilogue
    opIn.L = &L;
    opIn.dof = dof;
    opIn.K = Kin;

    ls.callD = wall_seconds();
    ls.suc = 0;
    hipLaunchKernelGGL(k_normsq_partial<double>, dim3(gv), dim3(256), 0, s, n, (const double*)R, part);
    hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, s, (int)gv, part, scal);
    double rr = 0.0;
    SV_HIP(hipMemcpyAsync(&rr, scal, sizeof(double), hipMemcpyDeviceToHost, s));
    SV_HIP(hipStreamSynchronize(s));
    ls.iNorm = std::sqrt(rr);
    ls.fNorm = ls.iNorm;
    const double eps = std::fmax(ls.absTol, ls.relTol * ls.iNorm);
    ls.itr = 0;
    ls.cM = 0;
    SV_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
    if (ls.iNorm <= ls.absTol) {
        ls.callD = 2.220446049250313e-16;
        ls.dB = 0.0;
        return SVFSI_OK;
    }
    SV_TRY(w.state.ensure(256));
    // inner-cycle buffers (allocated here so that no allocation happens
    // while a graph is being captured)
    {
        const size_t nbi = (n + (size_t)256 * VecE<IT>::E - 1) / ((size_t)256 * VecE<IT>::E);
        SV_TRY(w.basis32.ensure((size_t)(mi + 1) * n * sizeof(IT)));
        SV_TRY(w.partial.ensure(inner_partials(L.nNo, n, mi, nbi, gv) * sizeof(double)));
        SV_TRY(w.hess.ensure(gm_arrays_doubles(mi) * sizeof(double)));
    }
    hipLaunchKernelGGL(k_gm_init, dim3(1), dim3(64), 0, s, st, eps);
    hipLaunchKernelGGL(k_gm_init, dim3(1), dim3(64), 0, s, w.state.as<GmState>(), 0.0);
    PinnedPoll& poll = w.poll;
    SV_TRY(poll.ensure(256));
    GmState hs{};
    int outerMv = 0;

    // Polynomial inner preconditioner (dof 3, 4; SVFSI_INNER_POLY=0 keeps the
    // inner GMRES cycle): d Arnoldi steps of the FP32 inner operator B from
    // a seeded random start give the harmonic Ritz values, and every outer
    // step then applies the fixed GMRES(d) residual polynomial of B, one
    // SpMV per root and no inner dot products.
    std::vector<int> kinds;
    int polySpmv = 0;
    long long polySetup = 0;
    IT *P0 = nullptr, *P1 = nullptr, *Wb = nullptr;
    const double* coefD = nullptr;
    bool usePoly = polyWanted;
    if (usePoly) {
        // degree 25 when the inner matrix fits half the Infinity Cache, else
        // 30 (measured, ms per Newton step: 200k-tet cylinder 6.7 at 25
        // against 7.2 at 30; 3.2M-tet cylinder 227 at 30 against 243 at 25;
        // LELAS 20^3 3.32 against 3.34)
        const bool cached = (size_t)L.nnz * dof * dof * sizeof(IT) <= ((size_t)128 << 20);
        const int d = env_int("SVFSI_POLY_DEG", cached ? 25 : 30);
        if (d < 2 || d > kMaxSD) {
            set_error("SVFSI_POLY_DEG must be in 2..%d", kMaxSD);
            return SVFSI_ERR_ARG;
        }
        SV_TRY(w.polyVec.ensure(3 * n * sizeof(IT)));
        P0 = w.polyVec.as<IT>();
        P1 = P0 + n;
        Wb = P1 + n;
        SV_TRY(w.basis32.ensure((size_t)(std::max(mi, d) + 1) * n * sizeof(IT)));
        {
            const size_t nbi = (n + (size_t)256 * VecE<IT>::E - 1) / ((size_t)256 * VecE<IT>::E);
            SV_TRY(w.partial.ensure(inner_partials(L.nNo, n, std::max(mi, d), nbi, gv) * sizeof(double)));
            SV_TRY(w.hess.ensure(gm_arrays_doubles(std::max(mi, d)) * sizeof(double)));
        }
        SV_TRY(launch_rand_masked<IT>(n, 0x5f3759dfu, w.W.as<double>(), P0, s));
        SV_TRY(gmres_inner_dev<IT>(L, opIn, dof, P0, w.basis32, d, 0.0, nullptr, s));
        GmArrays Ad = gm_carve(w.hess.as<double>(), d);
        std::vector<double> Hu((size_t)(d + 1) * d);
        SV_HIP(hipMemcpyAsync(Hu.data(), Ad.Hu, Hu.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        GmState ins{};
        SV_HIP(hipMemcpyAsync(&ins, w.state.p, sizeof(GmState), hipMemcpyDeviceToHost, s));
        SV_HIP(hipStreamSynchronize(s));
        polySetup = ins.iters;
        std::vector<double> coef;
        const int nf = ins.iconv == d ? poly_factors(d, d + 1, Hu.data(), coef, kinds) : -1;
        if (nf > 0) {
            SV_TRY(w.polyCoef.ensure(coef.size() * sizeof(double)));
            SV_HIP(hipMemcpyAsync(w.polyCoef.p, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice,
                                  s));
            coefD = w.polyCoef.as<double>();
            for (int f = 0; f < nf; ++f) {
                const bool last = f == nf - 1;
                if (kinds[f] == 1) polySpmv += last ? 0 : 1;
                else polySpmv += 1 + ((!last && coef[2 * f + 1] != 0.0) ? 1 : 0);
            }
        } else {
            usePoly = false; // no usable roots: the inner GMRES cycle
        }
        hipLaunchKernelGGL(k_gm_init, dim3(1), dim3(64), 0, s, w.state.as<GmState>(), 0.0);
    }
    int amgSpmv = 0;
    if (useAmg) {
        if constexpr (std::is_same<IT, float>::value) {
            SV_TRY(w.polyVec.ensure(3 * n * sizeof(IT)));
            P0 = w.polyVec.as<IT>();
            SV_TRY(amg_setup(L, dof, K, Kin, s));
            std::vector<int> lv;
            amg_levels(L, lv);
            amgSpmv = lv.size() > 1 ? 4 : 0; // finest-level products per V-cycle
        }
    }
    // graph signature: every pointer and size baked into the step launches.
    // The polynomial steps are launched directly: their ~25 SpMV kernels of
    // ~8 us keep the queue full, and replaying them as graphs (keyed by the
    // factor sequence, bit for bit the same results) measured slower: 6.65
    // against 6.41 ms per Newton step on the 200k-tet cylinder, 3.07 against
    // 3.05 ms on the LELAS cube.
    const bool useGraphs = graphs_enabled() && !usePoly && !useAmg;
    if (useGraphs) {
        double tinv = tin;
        uintptr_t tbits;
        std::memcpy(&tbits, &tinv, sizeof(tbits));
        std::vector<uintptr_t> sig = {
            (uintptr_t)n, (uintptr_t)dof, (uintptr_t)mo, (uintptr_t)mi, tbits, (uintptr_t)K, (uintptr_t)Kin,
            (uintptr_t)Q, (uintptr_t)Z, (uintptr_t)t64, (uintptr_t)w.basis32.p, (uintptr_t)w.partial.p,
            (uintptr_t)w.hess.p, (uintptr_t)w.state.p, (uintptr_t)part, (uintptr_t)st, (uintptr_t)w.ohess.p,
            (uintptr_t)L.rowPtr.p, (uintptr_t)L.colPtr.p, (uintptr_t)fused_inner_enabled(), (uintptr_t)dinv,
            (uintptr_t)cgsMode, (uintptr_t)sizeof(IT), (uintptr_t)((long long)nbr > foldAt), (uintptr_t)fuseNorm};
        cpl_sig(L, sig);
        if (sig != w.graphs.sig) {
            w.graphs.clear();
            w.graphs.sig = sig;
        }
    }
    long long polySteps = 0;
    for (int l = 1; l <= ls.mItr; ++l) {
        ls.dB = ls.fNorm;
        ls.itr += 1;
        SV_TRY(launch_spmv(L, dof, K, x, Q, s));
        SV_TRY(cpl_apply(L, 0, dof, x, Q, s));
        hipLaunchKernelGGL(k_gm_resid<double>, dim3(gv), dim3(256), 0, s, n, (const double*)R, Q, part, st, A);
        hipLaunchKernelGGL(k_gm_div<double>, dim3(gv), dim3(256), 0, s, n, Q, A);
        outerMv += 1;
        int posted = 0;
        for (int i = 1; i <= mo; ++i) {
            // one outer step: inner cycle on z_i = M^-1 q_i, then the
            // FP64 product and the CGS2 Arnoldi step; the same launches
            // every cycle, so they replay as one graph per i
            auto step = [&, i](hipStream_t cs) -> int {
                IT* zi = Z + (size_t)(i - 1) * n;
                bool t64Done = false;
                if (useAmg) {
                    if constexpr (std::is_same<IT, float>::value) {
                        hipLaunchKernelGGL((k_cvt<double, IT>), dim3(gv), dim3(256), 0, cs, n,
                                           (const double*)(Q + (size_t)(i - 1) * n), P0);
                        SV_TRY(amg_apply(L, dof, P0, zi, cs));
                    }
                } else if (usePoly) {
                    // z_i = p(B) D^-1 q_i, factor by factor
                    launch_fg_start<IT>(L.nNo, dof, (const double*)(Q + (size_t)(i - 1) * n), dinv, P0, zi, cs);
                    IT* cur = P0;
                    IT* nxt = P1;
                    const int nf = (int)kinds.size();
                    for (int f = 0; f < nf; ++f) {
                        const bool last = f == nf - 1;
                        if (kinds[f] == 1) {
                            if (last) {
                                // with the FP64 copy for the outer product
                                hipLaunchKernelGGL(k_poly_last_cvt<IT>, dim3(gv), dim3(256), 0, cs, n,
                                                   (const IT*)cur, zi, t64, coefD, f);
                                t64Done = true;
                            } else {
                                SV_TRY(launch_poly<IT>(L, dof, 0, (const IT*)Kin, nullptr, cur, nullptr, nullptr, zi, nxt, coefD, f, cs,
                                                    &st->stop));
                                std::swap(cur, nxt);
                            }
                        } else {
                            SV_TRY(launch_poly<IT>(L, dof, 1, (const IT*)Kin, nullptr, cur, nullptr, Wb, zi, nullptr, coefD, f, cs,
                                                &st->stop));
                            if (!last) {
                                SV_TRY(launch_poly<IT>(L, dof, 2, (const IT*)Kin, nullptr, Wb, cur, nullptr, nullptr, nxt, coefD, f, cs,
                                                    &st->stop));
                                std::swap(cur, nxt);
                            }
                        }
                    }
                } else {
                    launch_fg_start<IT>(L.nNo, dof, (const double*)(Q + (size_t)(i - 1) * n), dinv, zi,
                                        (IT*)nullptr, cs);
                    SV_TRY(gmres_inner_dev<IT>(L, opIn, dof, zi, w.basis32, mi, tin, &st->stop, cs));
                }
                if (!t64Done)
                    hipLaunchKernelGGL((k_cvt<IT, double>), dim3(gv), dim3(256), 0, cs, n, (const IT*)zi, t64);
                SV_TRY(launch_spmv(L, dof, K, t64, Q + (size_t)i * n, cs, &st->stop));
                SV_TRY(cpl_apply(L, 0, dof, t64, Q + (size_t)i * n, cs, &st->stop));
                // outer Arnoldi: classical Gram-Schmidt, two passes (CGS2) by
                // default; SVFSI_OUTER_CGS=1 runs the reference's single pass
                for (int pp = 1; pp <= cgsPasses; ++pp) {
                    const int pass = cgsMode == 1 ? 0 : (cgsMode == 3 ? 1 : pp);
                    hipLaunchKernelGGL(k_gm_mdot<double>, dim3((unsigned)nbr, (unsigned)((i + 8) / 8)), dim3(256), 0,
                                       cs, n, i, (const double*)Q, part, (const GmState*)st);
                    if ((long long)nbr > foldAt) {
                        hipLaunchKernelGGL(k_gm_fold, dim3((unsigned)(i + 1)), dim3(256), 0, cs, (int)nbr, part,
                                           (const GmState*)st);
                        hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(256), 0, cs, i, mo, 1, (int)nbr,
                                           (const double*)part, st, A, pass);
                    } else {
                        hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(256), 0, cs, i, mo, (int)nbr, (int)nbr,
                                           (const double*)part, st, A, pass);
                    }
                    hipLaunchKernelGGL(k_gm_update<double>, dim3(gv), dim3(256), 0, cs, n, i, mo, Q,
                                       (const GmState*)st, A, cgsMode == 3 ? part : nullptr);
                    if (cgsMode == 3 && fuseNorm) {
                        hipLaunchKernelGGL(k_gm_normscale<double>, dim3(gv), dim3(256), 0, cs, n, i, mo, (int)gv,
                                           (const double*)part, Q, st, A);
                    } else if (cgsMode == 3) {
                        hipLaunchKernelGGL(k_gm_givens, dim3(1), dim3(256), 0, cs, i, mo, (int)gv, (int)gv,
                                           (const double*)part, st, A, 3);
                        hipLaunchKernelGGL(k_gm_scale<double>, dim3(gv), dim3(256), 0, cs, n, i, mo, Q,
                                           (const GmState*)st, A);
                    }
                }
                SV_HIP(hipGetLastError());
                return SVFSI_OK;
            };
            if (useGraphs)
                SV_TRY(graph_run(w.graphs, i, s, step));
            else
                SV_TRY(step(s));
            outerMv += 1;
            if (usePoly || useAmg) ++polySteps;
            // lagged poll: the state of step i-1 is read while step i runs
            const int q = posted & 1;
            SV_TRY(poll.post(q, st, sizeof(GmState), s));
            ++posted;
            if (posted >= 2) {
                SV_TRY(poll.wait(q ^ 1));
                std::memcpy(&hs, poll.h[q ^ 1], sizeof(GmState));
                if (hs.stop) break;
            }
        }
        {
            const size_t tri = (size_t)mo * (mo + 1) / 2;
            const size_t ldsB = (tri <= (size_t)kBackLds ? tri : 0) * sizeof(double);
            hipLaunchKernelGGL(k_gm_backsolve, dim3(1), dim3(64), ldsB, s, mo, st, A);
        }
        hipLaunchKernelGGL(k_fg_xupdate<IT>, dim3(gv), dim3(256), 0, s, n, x, (const IT*)Z, (const GmState*)st, A);
        SV_HIP(hipGetLastError());
        SV_HIP(hipMemcpyAsync(&hs, st, sizeof(GmState), hipMemcpyDeviceToHost, s));
        SV_HIP(hipStreamSynchronize(s));
        ls.itr += hs.iconv;
        ls.fNorm = hs.fNorm;
        if (hs.suc) {
            ls.suc = 1;
            break;
        }
    }
    {
        GmState in{};
        SV_HIP(hipMemcpyAsync(&in, w.state.p, sizeof(GmState), hipMemcpyDeviceToHost, s));
        SV_HIP(hipStreamSynchronize(s));
        // SpMVs: inner GMRES steps (or the polynomial's setup Arnoldi steps
        // plus its products), and the outer FP64 products (an outer step
        // launched after convergence exits at once but is counted)
        ls.cM = useAmg    ? (int)(polySteps * amgSpmv + outerMv)
                : usePoly ? (int)(polySetup + polySteps * polySpmv + outerMv)
                          : in.iters + outerMv;
    }
    SV_HIP(hipMemcpyAsync(R, x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    SV_HIP(hipStreamSynchronize(s));
    ls.callD = wall_seconds() - ls.callD;
    ls.dB = 10.0 * std::log(ls.fNorm / ls.dB);
    return SVFSI_OK;
}

int mixed_solve(Lhs& L, FSILS_subLsType& ls, int dof, const double* K, float* K32, double* R, hipStream_t s)
{
    return nested_solve<float>(L, ls, dof, K, K32, R, s);
}

int nested_solve_fp64(Lhs& L, FSILS_subLsType& ls, int dof, const double* K, double* R, hipStream_t s)
{
    return nested_solve<double>(L, ls, dof, K, nullptr, R, s);
}

} // namespace svfsi
