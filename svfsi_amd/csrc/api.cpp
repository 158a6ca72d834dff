// api.cpp -- extern "C" entry points of libsvfsi_amd (see include/svfsi_amd.h).
#include "internal.hpp"
#include "assemble.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace svfsi {

static thread_local std::string g_err;

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

Prof& prof()
{
    static Prof p;
    return p;
}

void prof_resolve()
{
    Prof& P = prof();
    for (auto& e : P.pending) {
        (void)hipEventSynchronize(e.b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e.a, e.b);
        P.count[e.k] += 1;
        P.ms[e.k] += ms;
        P.bytes[e.k] += e.bytes;
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    P.pending.clear();
}

int require_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        set_error("no HIP device: libsvfsi_amd has no CPU fallback");
        return SVFSI_ERR_NODEVICE;
    }
    return SVFSI_OK;
}

// TET4 rule exactly as the oracle / GETGIP+GETGNN compute it.
Tet4Rule make_tet4_rule()
{
    Tet4Rule r;
    const double s = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0;
    const double t = (5.0 - std::sqrt(5.0)) / 20.0;
    for (int g = 0; g < 4; ++g) {
        double xi[3] = {t, t, t};
        if (g < 3) xi[g] = s;
        r.w[g] = 1.0 / 24.0;
        r.N[g][0] = xi[0];
        r.N[g][1] = xi[1];
        r.N[g][2] = xi[2];
        r.N[g][3] = 1.0 - xi[0] - xi[1] - xi[2];
    }
    return r;
}

// GETGIP + GETGNN for TET10 (svFSI/NN.f:277-307, 674-715): the 15-point
// rule (quad 0) or the 4-point TET4 rule (quad 1), tabulated as
// w, N(10), Nxi(10,3) per Gauss point.
int make_tet10_rule(int quad, std::vector<double>& tab, std::vector<double>* xiOut)
{
    double w[15], xi[15][3];
    int nG;
    if (quad == 1) {
        const double s = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0;
        const double t = (5.0 - std::sqrt(5.0)) / 20.0;
        nG = 4;
        for (int g = 0; g < 4; ++g) {
            w[g] = 1.0 / 24.0;
            xi[g][0] = xi[g][1] = xi[g][2] = t;
            if (g < 3) xi[g][g] = s;
        }
    } else {
        nG = 15;
        w[0] = 0.030283678097089;
        for (int g = 1; g < 5; ++g) w[g] = 0.006026785714286;
        for (int g = 5; g < 9; ++g) w[g] = 0.011645249086029;
        for (int g = 9; g < 15; ++g) w[g] = 0.010949141561386;
        xi[0][0] = xi[0][1] = xi[0][2] = 0.25;
        // corner-type points: one coordinate t, the others s (last: all s)
        const double cs[2][2] = {{0.333333333333333, 0.0}, {0.090909090909091, 0.727272727272727}};
        for (int q = 0; q < 2; ++q)
            for (int g = 0; g < 4; ++g)
                for (int i = 0; i < 3; ++i) xi[1 + 4 * q + g][i] = (i == g) ? cs[q][1] : cs[q][0];
        const double s = 0.066550153573664, t = 0.433449846426336;
        const double e6[6][3] = {{s, s, t}, {s, t, s}, {s, t, t}, {t, t, s}, {t, s, t}, {t, s, s}};
        for (int g = 0; g < 6; ++g)
            for (int i = 0; i < 3; ++i) xi[9 + g][i] = e6[g][i];
    }
    tab.assign((size_t)nG * kTet10Rec, 0.0);
    for (int g = 0; g < nG; ++g) {
        double* r = tab.data() + (size_t)g * kTet10Rec;
        const double x = xi[g][0], y = xi[g][1], z = xi[g][2], s = 1.0 - x - y - z;
        r[0] = w[g];
        double* N = r + 1;
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
        const double d[10][3] = {{4.0 * x - 1.0, 0.0, 0.0},
                                 {0.0, 4.0 * y - 1.0, 0.0},
                                 {0.0, 0.0, 4.0 * z - 1.0},
                                 {1.0 - 4.0 * s, 1.0 - 4.0 * s, 1.0 - 4.0 * s},
                                 {4.0 * y, 4.0 * x, 0.0},
                                 {0.0, 4.0 * z, 4.0 * y},
                                 {4.0 * z, 0.0, 4.0 * x},
                                 {4.0 * (s - x), -4.0 * x, -4.0 * x},
                                 {-4.0 * y, 4.0 * (s - y), -4.0 * y},
                                 {-4.0 * z, -4.0 * z, 4.0 * (s - z)}};
        for (int a = 0; a < 10; ++a)
            for (int j = 0; j < 3; ++j) r[11 + a * 3 + j] = d[a][j];
    }
    if (xiOut) xiOut->assign(&xi[0][0], &xi[0][0] + (size_t)nG * 3);
    return nG;
}

int make_th_rule(std::vector<double>& tab)
{
    tab.assign((size_t)19 * kThRec, 0.0);
    for (int q = 0; q < 2; ++q) {
        std::vector<double> t10, xi;
        const int nG = make_tet10_rule(q, t10, &xi);
        for (int g = 0; g < nG; ++g) {
            double* r = tab.data() + (size_t)(q * 15 + g) * kThRec;
            for (int k = 0; k < kTet10Rec; ++k) r[k] = t10[(size_t)g * kTet10Rec + k];
            // the tet4 basis at the same point (GETGNN, svFSI/NN.f:655-671)
            const double* x = &xi[(size_t)g * 3];
            r[41] = x[0];
            r[42] = x[1];
            r[43] = x[2];
            r[44] = 1.0 - x[0] - x[1] - x[2];
        }
    }
    return 19;
}

Tri3Rule make_tri3_rule()
{
    Tri3Rule r;
    const double s = 2.0 / 3.0, t = 1.0 / 6.0;
    const double xi[3][2] = {{t, t}, {s, t}, {t, s}};
    for (int g = 0; g < 3; ++g) {
        r.w[g] = 1.0 / 6.0;
        r.N[g][0] = xi[g][0];
        r.N[g][1] = xi[g][1];
        r.N[g][2] = 1.0 - xi[g][0] - xi[g][1];
    }
    return r;
}

AsmParams to_asm(const svfsi_eq_params& p)
{
    AsmParams a;
    a.am = p.am; a.af = p.af; a.beta = p.beta; a.gam = p.gam; a.dt = p.dt;
    a.rho = p.rho; a.E = p.E; a.nu = p.nu;
    a.fb[0] = p.fb[0]; a.fb[1] = p.fb[1]; a.fb[2] = p.fb[2];
    a.mu = p.mu; a.bstab = p.bstab; a.C10 = p.C10; a.Kpen = p.Kpen;
    a.volType = p.volType; a.pad = 0;
    a.viscType = p.viscType == 0 ? SVFSI_VISC_CONST : p.viscType; a.pad2 = 0;
    a.mu_o = p.mu_o; a.lam = p.lam; a.va = p.visc_a; a.vn = p.visc_n;
    return a;
}

// Greedy colouring of items (elements or faces) so that no two items of a
// colour share a node; items are then listed colour by colour, ascending
// item id inside a colour.
static void color_items(int nItems, int nPer, const int* conn, int nNo, std::vector<int>& order,
                        std::vector<int>& off)
{
    std::vector<std::vector<uint64_t>> used(nNo);
    std::vector<int> col(nItems, 0);
    int nC = 0;
    for (int e = 0; e < nItems; ++e) {
        const int* c = conn + (size_t)e * nPer;
        int k = 0;
        for (;; ++k) {
            const int w = k >> 6;
            const uint64_t bit = 1ull << (k & 63);
            bool free_ = true;
            for (int a = 0; a < nPer && free_; ++a) {
                const auto& u = used[c[a]];
                if ((int)u.size() > w && (u[w] & bit)) free_ = false;
            }
            if (free_) break;
        }
        col[e] = k;
        const int w = k >> 6;
        for (int a = 0; a < nPer; ++a) {
            auto& u = used[c[a]];
            if ((int)u.size() <= w) u.resize(w + 1, 0);
            u[w] |= 1ull << (k & 63);
        }
        if (k + 1 > nC) nC = k + 1;
    }
    off.assign(nC + 1, 0);
    for (int e = 0; e < nItems; ++e) off[col[e] + 1]++;
    for (int c = 0; c < nC; ++c) off[c + 1] += off[c];
    order.resize(nItems);
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int e = 0; e < nItems; ++e) order[pos[col[e]]++] = e;
}

static int find_block(const Lhs& L, int row, int col)
{
    const int* cp = L.hColPtr.data();
    int lo = L.hRowPtr[row], hi = L.hRowPtr[row + 1];
    const int* it = std::lower_bound(cp + lo, cp + hi, col);
    if (it != cp + hi && *it == col) return (int)(it - cp);
    for (int k = lo; k < hi; ++k)
        if (cp[k] == col) return k; // unsorted rows
    return -1;
}

template <typename T>
static int upload(DBuf& b, const std::vector<T>& v)
{
    SV_TRY(b.ensure(v.size() * sizeof(T)));
    if (!v.empty()) SV_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return SVFSI_OK;
}

static int solve_impl(Lhs& L, FSILS_lsType* ls, int dof, double* R, double* Val, int prec,
                      const int* incL, const double* res, hipStream_t s, int mode)
{
    SV_ARG(ls, "ls is NULL");
    SV_ARG(dof >= 2 && dof <= 4, "dof %d not supported (2..4)", dof);
    if (prec != PRECOND_FSILS && prec != PRECOND_RCS) {
        set_error("preconditioner %d not supported (PRECOND_FSILS=%d, PRECOND_RCS=%d)", prec, PRECOND_FSILS,
                  PRECOND_RCS);
        return SVFSI_ERR_NOTSUP;
    }
    if (ls->LS_type != LS_TYPE_GMRES && ls->LS_type != LS_TYPE_CG && ls->LS_type != LS_TYPE_NS &&
        ls->LS_type != LS_TYPE_BICGS) {
        set_error("LS_type %d not supported (BICGS 795, NS 796, GMRES 797, CG 798)", ls->LS_type);
        return SVFSI_ERR_NOTSUP;
    }
    if (ls->LS_type == LS_TYPE_NS && (mode != 0 || dof != 4)) {
        set_error("the NS solver is FP64 only and needs dof 4 (velocity 3 + pressure)");
        return SVFSI_ERR_NOTSUP;
    }
    // coupled Neumann faces (SOLVE.f:66-86): the rank-one terms enter GMRES
    // and NS products (ADDBCMUL); CGRADV has none, so CG solves without them
    struct CplOff {
        CplSet& c;
        ~CplOff() { c.active = false; }
    } cplOff{L.cpl};
    SV_TRY(cpl_setup(L, dof, incL, res, s));
    // BICGSV has no coupled terms either
    if (ls->LS_type == LS_TYPE_CG || ls->LS_type == LS_TYPE_BICGS) L.cpl.active = false;
    if (L.cpl.active && prec == PRECOND_RCS) {
        set_error("coupled Neumann faces need PRECOND_FSILS: PRECONDRCS leaves the scaled face vectors unset "
                  "(PRECOND.f:356-364)");
        return SVFSI_ERR_NOTSUP;
    }
    const size_t n = (size_t)L.nNo * dof;
    Work& w = L.ws;
    SV_TRY(w.W.ensure(n * sizeof(double)));
    // SVFSI_INNER_FP64=1: the mixed entry points run the same nested solver
    // with its inner preconditioner in FP64 (a comparison of precisions)
    if (mode == 1) {
        const char* e = getenv("SVFSI_INNER_FP64");
        if (e && e[0] == '1') mode = 2;
    }
    float* v32 = nullptr;
    if (mode == 1) {
        SV_TRY(w.val32.ensure((size_t)L.nnz * dof * dof * sizeof(float)));
        v32 = w.val32.as<float>();
    }
    // w.W ends up holding the factor that unscales the solution (SOLVE.f:134):
    // W of PRECONDDIAG, or W2 (the column factor) of PRECONDRCS
    if (prec == PRECOND_RCS) {
        SV_TRY(w.W1.ensure(n * sizeof(double)));
        SV_TRY(precond_rcs(L, dof, Val, R, w.W1.as<double>(), w.W.as<double>(), v32, incL, s));
    } else {
        SV_TRY(precond_diag(L, dof, Val, R, w.W.as<double>(), v32, incL, s));
        SV_TRY(cpl_scale(L, w.W.as<double>(), s));
    }
    int rc;
    if (mode == 1) rc = mixed_solve(L, ls->RI, dof, Val, v32, R, s);
    else if (mode == 2) rc = nested_solve_fp64(L, ls->RI, dof, Val, R, s);
    else if (ls->LS_type == LS_TYPE_NS) rc = ns_solve(L, *ls, dof, Val, R, s);
    else if (ls->LS_type == LS_TYPE_GMRES) rc = gmres_solve(L, ls->RI, dof, Val, R, s, false, L.cpl.active ? 1 : 0);
    else if (ls->LS_type == LS_TYPE_BICGS) rc = bicgs_solve(L, ls->RI, dof, Val, R, s);
    else rc = cg_solve(L, ls->RI, dof, Val, R, s);
    if (rc != SVFSI_OK) return rc;
    SV_TRY(launch_scale_vec(n, w.W.as<double>(), R, s));
    SV_HIP(hipStreamSynchronize(s));
    return SVFSI_OK;
}

} // namespace svfsi

using namespace svfsi;

extern "C" {

const char* svfsi_last_error(void) { return g_err.c_str(); }

int svfsi_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int svfsi_dev_malloc(void** dptr, size_t bytes)
{
    SV_ARG(dptr, "dptr is NULL");
    SV_TRY(require_device());
    *dptr = nullptr;
    SV_HIP(hipMalloc(dptr, bytes ? bytes : 8));
    return SVFSI_OK;
}

int svfsi_dev_free(void* dptr)
{
    if (dptr) SV_HIP(hipFree(dptr));
    return SVFSI_OK;
}

int svfsi_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    SV_ARG(dst && src, "NULL argument");
    SV_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SVFSI_OK;
}

int svfsi_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    SV_ARG(dst && src, "NULL argument");
    SV_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SVFSI_OK;
}

int fsils_gpu_ls_create(FSILS_lsType* ls, int LS_type, const double* relTol, const double* absTol,
                        const int* maxItr, const int* dimKry, const double* relTolIn,
                        const double* absTolIn, const int* maxItrIn)
{
    SV_ARG(ls, "ls is NULL");
    if (ls->foC) {
        set_error("FSILS: LS is not free (LS.f:60-63)");
        return SVFSI_ERR_STATE;
    }
    std::memset(ls, 0, sizeof(*ls));
    ls->foC = 1;
    ls->LS_type = LS_type;
    switch (LS_type) {
    case LS_TYPE_NS:
        ls->RI.relTol = 0.4; ls->GM.relTol = 1e-2; ls->CG.relTol = 0.2;
        ls->RI.mItr = 10; ls->GM.mItr = 2; ls->CG.mItr = 500;
        ls->GM.sD = 100; ls->RI.sD = 100;
        break;
    case LS_TYPE_GMRES:
        ls->RI.relTol = 0.1; ls->RI.mItr = 4; ls->RI.sD = 250;
        break;
    case LS_TYPE_CG:
        ls->RI.relTol = 1e-2; ls->RI.mItr = 1000;
        break;
    case LS_TYPE_BICGS:
        ls->RI.relTol = 1e-2; ls->RI.mItr = 500;
        break;
    default:
        ls->foC = 0;
        set_error("FSILS: LS_TYPE is not defined (LS.f:89-91)");
        return SVFSI_ERR_ARG;
    }
    ls->RI.absTol = 1e-10;
    ls->GM.absTol = 1e-10;
    ls->CG.absTol = 1e-10;
    if (relTol) ls->RI.relTol = *relTol;
    if (absTol) ls->RI.absTol = *absTol;
    if (maxItr) ls->RI.mItr = *maxItr;
    if (dimKry) { ls->RI.sD = *dimKry; ls->GM.sD = *dimKry; }
    if (relTolIn) { ls->GM.relTol = relTolIn[0]; ls->CG.relTol = relTolIn[1]; }
    if (absTolIn) { ls->GM.absTol = absTolIn[0]; ls->CG.absTol = absTolIn[1]; }
    if (maxItrIn) { ls->GM.mItr = maxItrIn[0]; ls->CG.mItr = maxItrIn[1]; }
    return SVFSI_OK;
}

int fsils_gpu_lhs_create(void** lhs, int gnNo, int nNo, int nnz, const int* gNodes, const int* rowPtr,
                         const int* colPtr, int nFaces)
{
    (void)gNodes;
    SV_ARG(lhs, "lhs is NULL");
    SV_ARG(nNo > 0 && nnz > 0, "nNo (%d) and nnz (%d) must be positive", nNo, nnz);
    SV_ARG(rowPtr && colPtr, "rowPtr/colPtr NULL");
    SV_ARG(nFaces >= 0, "nFaces < 0");
    SV_ARG(rowPtr[0] == 1 && rowPtr[nNo] == nnz + 1, "rowPtr must be 1-based with rowPtr(nNo+1) = nnz+1");
    SV_TRY(require_device());
    Lhs* L = new Lhs();
    L->gnNo = gnNo;
    L->nNo = nNo;
    L->nnz = nnz;
    L->nFaces = nFaces;
    L->faces.resize(nFaces);
    L->hRowPtr.resize(nNo + 1);
    L->hColPtr.resize(nnz);
    for (int i = 0; i <= nNo; ++i) L->hRowPtr[i] = rowPtr[i] - 1;
    for (int i = 0; i < nNo; ++i)
        if (L->hRowPtr[i + 1] < L->hRowPtr[i]) {
            delete L;
            set_error("rowPtr not monotone at row %d", i + 1);
            return SVFSI_ERR_ARG;
        }
    for (int k = 0; k < nnz; ++k) {
        L->hColPtr[k] = colPtr[k] - 1;
        if (colPtr[k] < 1 || colPtr[k] > nNo) {
            delete L;
            set_error("colPtr(%d) = %d out of range", k + 1, colPtr[k]);
            return SVFSI_ERR_ARG;
        }
    }
    std::vector<int> diag(nNo), brow(nnz), trans(nnz);
    for (int i = 0; i < nNo; ++i) {
        diag[i] = find_block(*L, i, i);
        if (diag[i] < 0) {
            delete L;
            set_error("row %d has no diagonal block", i + 1);
            return SVFSI_ERR_ARG;
        }
        for (int k = L->hRowPtr[i]; k < L->hRowPtr[i + 1]; ++k) brow[k] = i;
    }
    L->symmetric = true;
    for (int k = 0; k < nnz; ++k) {
        trans[k] = find_block(*L, L->hColPtr[k], brow[k]);
        if (trans[k] < 0) L->symmetric = false;
    }
    int rc = SVFSI_OK;
    if ((rc = upload(L->rowPtr, L->hRowPtr)) || (rc = upload(L->colPtr, L->hColPtr)) ||
        (rc = upload(L->diagPtr, diag)) || (rc = upload(L->blockRow, brow)) || (rc = upload(L->trans, trans)) ||
        (rc = build_row_tiles(*L))) {
        delete L;
        return rc;
    }
    *lhs = L;
    return SVFSI_OK;
}

int fsils_gpu_lhs_free(void* lhs)
{
    SV_ARG(lhs, "lhs is NULL");
    prof_resolve();
    delete static_cast<Lhs*>(lhs);
    return SVFSI_OK;
}

int fsils_gpu_bc_create(void* lhs, int faIn, int nNo, int dof, int BC_type, const int* gNodes,
                        const double* Val)
{
    SV_ARG(lhs, "lhs is NULL");
    Lhs& L = *static_cast<Lhs*>(lhs);
    SV_ARG(faIn >= 1 && faIn <= L.nFaces, "faIn %d out of range 1..%d (BC.f:63-66)", faIn, L.nFaces);
    SV_ARG(BC_type == BC_TYPE_Dir || BC_type == BC_TYPE_Neu, "undefined BC_type %d (BC.f:74-77)", BC_type);
    SV_ARG(nNo >= 0 && dof >= 1, "bad nNo/dof");
    Face& F = L.faces[faIn - 1];
    if (F.foC) {
        set_error("FSILS: face is not free (BC.f:67-70)");
        return SVFSI_ERR_STATE;
    }
    F.foC = true;
    F.nNo = nNo;
    F.dof = dof;
    F.bGrp = BC_type;
    F.glob.resize(nNo);
    F.val.assign((size_t)nNo * dof, 0.0);
    for (int a = 0; a < nNo; ++a) {
        const int g = gNodes[a] - 1;
        if (g < 0 || g >= L.nNo) {
            F = Face();
            set_error("gNodes(%d) = %d out of range", a + 1, gNodes[a]);
            return SVFSI_ERR_ARG;
        }
        F.glob[a] = g;
    }
    if (Val)
        for (size_t q = 0; q < F.val.size(); ++q) F.val[q] = Val[q];
    L.facesVersion++;
    return SVFSI_OK;
}

int fsils_gpu_bc_free(void* lhs, int faIn)
{
    SV_ARG(lhs, "lhs is NULL");
    Lhs& L = *static_cast<Lhs*>(lhs);
    SV_ARG(faIn >= 1 && faIn <= L.nFaces, "faIn out of range");
    if (!L.faces[faIn - 1].foC) {
        set_error("FSILS: Cannot free a face that is not created yet (BC.f:200-203)");
        return SVFSI_ERR_STATE;
    }
    L.faces[faIn - 1] = Face();
    L.facesVersion++;
    return SVFSI_OK;
}

int fsils_gpu_solve_dev(void* lhs, FSILS_lsType* ls, int dof, double* dRi, double* dVal, int prec,
                        const int* incL, const double* res, void* stream)
{
    SV_ARG(lhs && dRi && dVal, "NULL argument");
    SV_TRY(require_device());
    return solve_impl(*static_cast<Lhs*>(lhs), ls, dof, dRi, dVal, prec, incL, res,
                      static_cast<hipStream_t>(stream), 0);
}

int fsils_gpu_solve_mixed_dev(void* lhs, FSILS_lsType* ls, int dof, double* dRi, double* dVal, int prec,
                              const int* incL, const double* res, void* stream)
{
    SV_ARG(lhs && dRi && dVal, "NULL argument");
    SV_TRY(require_device());
    return solve_impl(*static_cast<Lhs*>(lhs), ls, dof, dRi, dVal, prec, incL, res,
                      static_cast<hipStream_t>(stream), 1);
}

int fsils_gpu_solve(void* lhs, FSILS_lsType* ls, int dof, double* Ri, double* Val, int prec,
                    const int* incL, const double* res)
{
    SV_ARG(lhs && Ri && Val, "NULL argument");
    SV_TRY(require_device());
    Lhs& L = *static_cast<Lhs*>(lhs);
    const size_t nR = (size_t)L.nNo * dof, nV = (size_t)L.nnz * dof * dof;
    DBuf dR, dV;
    SV_TRY(dR.ensure(nR * sizeof(double)));
    SV_TRY(dV.ensure(nV * sizeof(double)));
    SV_HIP(hipMemcpy(dR.p, Ri, nR * sizeof(double), hipMemcpyHostToDevice));
    SV_HIP(hipMemcpy(dV.p, Val, nV * sizeof(double), hipMemcpyHostToDevice));
    SV_TRY(solve_impl(L, ls, dof, dR.as<double>(), dV.as<double>(), prec, incL, res, nullptr, 0));
    SV_HIP(hipMemcpy(Ri, dR.p, nR * sizeof(double), hipMemcpyDeviceToHost));
    SV_HIP(hipMemcpy(Val, dV.p, nV * sizeof(double), hipMemcpyDeviceToHost));
    return SVFSI_OK;
}

static int solve_t_impl(void* lhs, FSILS_lsType* ls, int dof, const double* dVal, const double* drhs,
                        double* dlam, const int* incL, void* stream, int mode)
{
    SV_ARG(lhs && ls && dVal && drhs && dlam, "NULL argument");
    SV_TRY(require_device());
    Lhs& L = *static_cast<Lhs*>(lhs);
    hipStream_t s = static_cast<hipStream_t>(stream);
    SV_ARG(dof >= 2 && dof <= 4, "dof %d not supported (2..4)", dof);
    SV_ARG(L.symmetric, "pattern is not structurally symmetric: no transposed solve");
    const size_t n = (size_t)L.nNo * dof, nV = (size_t)L.nnz * dof * dof;
    Work& w = L.ws;
    SV_TRY(w.valT.ensure(nV * sizeof(double)));
    // K^T on the same pattern; its diagonal (hence W and the Dirichlet
    // mask) is that of K, so the scaled system is (W K W)^T
    SV_TRY(launch_transpose_blocks(L, dof, dVal, w.valT.as<double>(), s));
    SV_HIP(hipMemcpyAsync(dlam, drhs, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    return solve_impl(L, ls, dof, dlam, w.valT.as<double>(), PRECOND_FSILS, incL, nullptr, s, mode);
}

int fsils_gpu_solve_t_dev(void* lhs, FSILS_lsType* ls, int dof, const double* dVal, const double* drhs,
                          double* dlam, const int* incL, void* stream)
{
    return solve_t_impl(lhs, ls, dof, dVal, drhs, dlam, incL, stream, 0);
}

int fsils_gpu_solve_t_mixed_dev(void* lhs, FSILS_lsType* ls, int dof, const double* dVal, const double* drhs,
                                double* dlam, const int* incL, void* stream)
{
    return solve_t_impl(lhs, ls, dof, dVal, drhs, dlam, incL, stream, 1);
}

int svfsi_spmv_dev(void* lhs, int dof, const double* dK, const double* dU, double* dKU, void* stream)
{
    SV_ARG(lhs && dK && dU && dKU, "NULL argument");
    SV_TRY(require_device());
    return launch_spmv(*static_cast<Lhs*>(lhs), dof, dK, dU, dKU, static_cast<hipStream_t>(stream));
}

int svfsi_spmv_f32_dev(void* lhs, int dof, const float* dK, const float* dU, float* dKU, void* stream)
{
    SV_ARG(lhs && dK && dU && dKU, "NULL argument");
    SV_TRY(require_device());
    return launch_spmv_f32(*static_cast<Lhs*>(lhs), dof, dK, dU, dKU, static_cast<hipStream_t>(stream));
}

int svfsi_spmv_t_dev(void* lhs, int dof, const double* dK, const double* dU, double* dKU, void* stream)
{
    SV_ARG(lhs && dK && dU && dKU, "NULL argument");
    SV_TRY(require_device());
    return launch_spmv_t(*static_cast<Lhs*>(lhs), dof, dK, dU, dKU, static_cast<hipStream_t>(stream));
}

int svfsi_amg_aggregate(int nNo, const int* rowPtr, const int* colPtr, int cap, int* agg)
{
    SV_ARG(nNo >= 0 && rowPtr && colPtr && agg && cap >= 2, "bad argument");
    return guarded([&]() -> int {
        for (int i = 0; i < nNo; ++i)
            if (rowPtr[i + 1] < rowPtr[i]) {
                set_error("rowPtr is not monotone at row %d", i);
                return SVFSI_ERR_ARG;
            }
        for (int k = 0; k < rowPtr[nNo]; ++k)
            if (colPtr[k] < 0 || colPtr[k] >= nNo) {
                set_error("colPtr[%d] = %d outside 0..%d", k, colPtr[k], nNo - 1);
                return SVFSI_ERR_ARG;
            }
        std::vector<int> a;
        const int na = amg_aggregate(nNo, rowPtr, colPtr, cap, a);
        std::copy(a.begin(), a.end(), agg);
        return na;
    });
}

int svfsi_amg_levels(void* lhs, int* nodes, int maxLevels)
{
    SV_ARG(lhs && nodes && maxLevels > 0, "bad argument");
    std::vector<int> lv;
    const int n = amg_levels(*static_cast<Lhs*>(lhs), lv);
    for (int i = 0; i < n && i < maxLevels; ++i) nodes[i] = lv[i];
    return n;
}

int svfsi_resid_dd_dev(void* lhs, int dof, const double* dK, const double* dU, const double* dB, double* dR,
                       int transpose, void* stream)
{
    SV_ARG(lhs && dK && dU && dB && dR, "NULL argument");
    SV_ARG(dR != dU, "dR must not alias dU");
    SV_TRY(require_device());
    return launch_resid_dd(*static_cast<Lhs*>(lhs), dof, dK, dU, dB, dR, transpose != 0,
                           static_cast<hipStream_t>(stream));
}

int svfsi_dot_dev(int n, const double* dU, const double* dV, double* out, void* stream)
{
    SV_ARG(dU && dV && out && n >= 0, "bad argument");
    SV_TRY(require_device());
    hipStream_t s = static_cast<hipStream_t>(stream);
    DBuf part;
    SV_TRY(part.ensure(520 * sizeof(double)));
    SV_TRY(launch_dot(n, dU, dV, part.as<double>(), part.as<double>() + 512, s));
    SV_HIP(hipMemcpyAsync(out, part.as<double>() + 512, sizeof(double), hipMemcpyDeviceToHost, s));
    SV_HIP(hipStreamSynchronize(s));
    return SVFSI_OK;
}

int svfsi_precond_diag_dev(void* lhs, int dof, double* dVal, double* dR, double* dW, const int* incL,
                           void* stream)
{
    SV_ARG(lhs && dVal && dR && dW, "NULL argument");
    SV_ARG(dof >= 2 && dof <= 4, "dof %d not supported (2..4)", dof);
    SV_TRY(require_device());
    hipStream_t s = static_cast<hipStream_t>(stream);
    SV_TRY(precond_diag(*static_cast<Lhs*>(lhs), dof, dVal, dR, dW, nullptr, incL, s));
    SV_HIP(hipStreamSynchronize(s));
    return SVFSI_OK;
}

int svfsi_precond_rcs_dev(void* lhs, int dof, double* dVal, double* dR, double* dW1, double* dW2,
                          const int* incL, void* stream)
{
    SV_ARG(lhs && dVal && dR && dW1 && dW2, "NULL argument");
    SV_ARG(dof >= 2 && dof <= 4, "dof %d not supported (2..4)", dof);
    SV_TRY(require_device());
    hipStream_t s = static_cast<hipStream_t>(stream);
    SV_TRY(precond_rcs(*static_cast<Lhs*>(lhs), dof, dVal, dR, dW1, dW2, nullptr, incL, s));
    SV_HIP(hipStreamSynchronize(s));
    return SVFSI_OK;
}

int svfsi_mesh_create(void** mesh, void* lhs, int nNo, int nEl, int eNoN, const int* IEN, const double* x)
{
    SV_ARG(mesh && lhs && IEN && x, "NULL argument");
    SV_ARG(eNoN == 4 || eNoN == 10, "eNoN %d not supported (4 or 10)", eNoN);
    SV_TRY(require_device());
    Lhs& L = *static_cast<Lhs*>(lhs);
    SV_ARG(nNo == L.nNo, "mesh nNo %d != lhs nNo %d", nNo, L.nNo);
    SV_ARG(nEl > 0, "nEl must be positive");
    std::vector<int> ien((size_t)nEl * eNoN);
    for (size_t q = 0; q < ien.size(); ++q) {
        ien[q] = IEN[q] - 1;
        SV_ARG(ien[q] >= 0 && ien[q] < nNo, "IEN entry %zu = %d out of range", q + 1, IEN[q]);
    }
    Mesh* M = new Mesh();
    M->lhs = &L;
    M->nNo = nNo;
    M->nEl = nEl;
    M->eNoN = eNoN;
    std::vector<int> order;
    color_items(nEl, eNoN, ien.data(), nNo, order, M->colorOff);
    std::vector<int> sien((size_t)nEl * eNoN), emap((size_t)nEl * eNoN * eNoN);
    for (int p = 0; p < nEl; ++p) {
        const int e = order[p];
        for (int a = 0; a < eNoN; ++a) sien[(size_t)p * eNoN + a] = ien[(size_t)e * eNoN + a];
        for (int a = 0; a < eNoN; ++a)
            for (int b = 0; b < eNoN; ++b) {
                const int k = find_block(L, ien[(size_t)e * eNoN + a], ien[(size_t)e * eNoN + b]);
                if (k < 0) {
                    delete M;
                    set_error("element %d couples nodes (%d,%d) absent from the pattern", e + 1,
                              ien[(size_t)e * eNoN + a] + 1, ien[(size_t)e * eNoN + b] + 1);
                    return SVFSI_ERR_ARG;
                }
                emap[((size_t)p * eNoN + a) * eNoN + b] = k;
            }
    }
    // gather lists, filled in element-id order
    const int NN = eNoN * eNoN;
    std::vector<int> slot(nEl);
    for (int p = 0; p < nEl; ++p) slot[order[p]] = p;
    std::vector<int> gptr((size_t)L.nnz + 1, 0), rptr((size_t)nNo + 1, 0);
    for (int k : emap) gptr[(size_t)k + 1]++;
    for (int nd : sien) rptr[(size_t)nd + 1]++;
    for (size_t j = 0; j < (size_t)L.nnz; ++j) gptr[j + 1] += gptr[j];
    for (size_t i = 0; i < (size_t)nNo; ++i) rptr[i + 1] += rptr[i];
    std::vector<int> gslot(emap.size()), rslot(sien.size());
    {
        std::vector<int> gpos(gptr.begin(), gptr.end() - 1), rpos(rptr.begin(), rptr.end() - 1);
        for (int e = 0; e < nEl; ++e) {
            const int p = slot[e];
            for (int ab = 0; ab < NN; ++ab) gslot[(size_t)p * NN + ab] = gpos[emap[(size_t)p * NN + ab]]++;
            for (int a = 0; a < eNoN; ++a) rslot[(size_t)p * eNoN + a] = rpos[sien[(size_t)p * eNoN + a]]++;
        }
    }
    std::vector<double> xv(x, x + (size_t)3 * nNo);
    // mid-edge nodes (local 5..10 of a tet10), whose pressure rows
    // THOOD_ValRC resets in a Taylor-Hood FLUID assembly (FS.f:383-391)
    std::vector<int> mids;
    if (eNoN == 10) {
        std::vector<char> isMid(nNo, 0);
        for (int e = 0; e < nEl; ++e)
            for (int a = 4; a < 10; ++a) isMid[ien[(size_t)e * 10 + a]] = 1;
        for (int n = 0; n < nNo; ++n)
            if (isMid[n]) mids.push_back(n);
    }
    M->nMid = (int)mids.size();
    int rc = SVFSI_OK;
    if ((rc = upload(M->ien, sien)) || (rc = upload(M->elemId, order)) || (rc = upload(M->emap, emap)) ||
        (rc = upload(M->x, xv)) || (rc = upload(M->gPtr, gptr)) || (rc = upload(M->gSlot, gslot)) ||
        (rc = upload(M->rPtr, rptr)) || (rc = upload(M->rSlot, rslot)) ||
        (!mids.empty() && (rc = upload(M->midNodes, mids)))) {
        delete M;
        return rc;
    }
    *mesh = M;
    return SVFSI_OK;
}

int svfsi_mesh_free(void* mesh)
{
    SV_ARG(mesh, "mesh is NULL");
    delete static_cast<Mesh*>(mesh);
    return SVFSI_OK;
}

int svfsi_mesh_ncolors(void* mesh)
{
    if (!mesh) return SVFSI_ERR_ARG;
    return (int)static_cast<Mesh*>(mesh)->colorOff.size() - 1;
}

int svfsi_mesh_add_face(void* mesh, int nFe, const int* fIEN, const int* fEl, int* faceId)
{
    SV_ARG(mesh && fIEN && fEl && faceId, "NULL argument");
    Mesh& M = *static_cast<Mesh*>(mesh);
    SV_ARG(nFe >= 0, "nFe < 0");
    std::vector<int> fien((size_t)nFe * 3), fel(nFe), inner(nFe), fmap((size_t)nFe * 9);
    std::vector<int> hien((size_t)M.nEl * M.eNoN), elemId(M.nEl);
    SV_HIP(hipMemcpy(hien.data(), M.ien.p, hien.size() * sizeof(int), hipMemcpyDeviceToHost));
    SV_HIP(hipMemcpy(elemId.data(), M.elemId.p, elemId.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> slot(M.nEl);
    for (int p = 0; p < M.nEl; ++p) slot[elemId[p]] = p;
    for (int f = 0; f < nFe; ++f) {
        for (int a = 0; a < 3; ++a) {
            fien[(size_t)f * 3 + a] = fIEN[(size_t)f * 3 + a] - 1;
            SV_ARG(fien[(size_t)f * 3 + a] >= 0 && fien[(size_t)f * 3 + a] < M.nNo, "face node out of range");
        }
        fel[f] = fEl[f] - 1;
        SV_ARG(fel[f] >= 0 && fel[f] < M.nEl, "face element out of range");
        const int* en = hien.data() + (size_t)slot[fel[f]] * M.eNoN;
        int in = -1, found = 0;
        for (int b = 0; b < M.eNoN; ++b) {
            const int nb = en[b];
            const bool onFace = nb == fien[f * 3] || nb == fien[f * 3 + 1] || nb == fien[f * 3 + 2];
            if (onFace) found++;
            else if (in < 0) in = nb;
        }
        SV_ARG(found == 3 && in >= 0, "face %d does not match element %d (GNNB, NN.f:1878-1905)", f + 1, fEl[f]);
        inner[f] = in;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                const int k = find_block(*M.lhs, fien[f * 3 + a], fien[f * 3 + b]);
                SV_ARG(k >= 0, "face block absent from pattern");
                fmap[(size_t)f * 9 + a * 3 + b] = k;
            }
    }
    auto* F = new Mesh::FaceSet();
    F->nFe = nFe;
    std::vector<int> order;
    color_items(nFe, 3, fien.data(), M.nNo, order, F->colorOff);
    int rc = SVFSI_OK;
    if ((rc = upload(F->fien, fien)) || (rc = upload(F->fel, fel)) || (rc = upload(F->inner, inner)) ||
        (rc = upload(F->fmap, fmap)) || (rc = upload(F->order, order))) {
        delete F;
        return rc;
    }
    M.faceSets.push_back(F);
    *faceId = (int)M.faceSets.size() - 1;
    return SVFSI_OK;
}

int svfsi_zero_dev(void* mesh, int dof, double* dVal, double* dR, void* stream)
{
    SV_ARG(mesh && dVal && dR, "NULL argument");
    SV_TRY(require_device());
    Mesh& M = *static_cast<Mesh*>(mesh);
    hipStream_t s = static_cast<hipStream_t>(stream);
    SV_HIP(hipMemsetAsync(dVal, 0, (size_t)M.lhs->nnz * dof * dof * sizeof(double), s));
    SV_HIP(hipMemsetAsync(dR, 0, (size_t)M.nNo * dof * sizeof(double), s));
    return SVFSI_OK;
}

int svfsi_assemble_dev(void* mesh, int phys, const svfsi_eq_params* p, int tDof, const double* dAg,
                       const double* dYg, const double* dDg, const double* dprop_el, int quad,
                       double* dVal, double* dR, void* stream)
{
    SV_ARG(mesh && p && dVal && dR, "NULL argument");
    SV_TRY(require_device());
    Mesh& M = *static_cast<Mesh*>(mesh);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (phys == SVFSI_PHYS_LELAS) {
        SV_ARG(M.eNoN == 4, "LELAS assembly supports tet4 only");
        SV_ARG(tDof >= 3 && dAg && dDg, "LELAS needs Ag, Dg with tDof >= 3");
    } else if (phys == SVFSI_PHYS_FLUID) {
        SV_ARG(tDof >= 4 && dAg && dYg, "FLUID needs Ag, Yg with tDof >= 4");
        const int vt = p->viscType;
        SV_ARG(vt == 0 || vt == SVFSI_VISC_CONST || vt == SVFSI_VISC_CY || vt == SVFSI_VISC_CASS,
               "viscType %d: SVFSI_VISC_CONST (698), SVFSI_VISC_CY (697) or SVFSI_VISC_CASS (696)", vt);
        SV_ARG(vt != SVFSI_VISC_CY || p->visc_a > 0.0, "Carreau-Yasuda needs a > 0");
        if (M.eNoN == 10) {
            if (quad != SVFSI_FLUID_P2P1) {
                set_error("FLUID on tet10: pass quad = SVFSI_FLUID_P2P1 (Taylor-Hood, nFs = 2); the VMS tet10 "
                          "element (nFs = 1) is not built");
                return SVFSI_ERR_NOTSUP;
            }
            return launch_asm_fluid_th(M, to_asm(*p), tDof, dAg, dYg, dprop_el, dVal, dR, s);
        }
        SV_ARG(quad == 0, "FLUID on tet4: quad must be 0 (SVFSI_FLUID_P2P1 needs a tet10 mesh)");
    } else if (phys == SVFSI_PHYS_STRUCT) {
        SV_ARG(M.eNoN == 10, "STRUCT assembly supports tet10 only");
        SV_ARG(tDof >= 3 && dAg && dDg, "STRUCT needs Ag, Dg with tDof >= 3");
        const int kern = quad & 0x300;
        SV_ARG((quad & ~0x301) == 0, "quad must be 0 (15-point) or 1 (4-point), ORed with SVFSI_STRUCT_*");
        quad &= 1;
        if (!M.rule10nG[quad]) {
            std::vector<double> tab;
            const int nG = make_tet10_rule(quad, tab);
            SV_TRY(upload(M.rule10[quad], tab));
            M.rule10nG[quad] = nG;
        }
        return launch_asm_struct(M, to_asm(*p), M.rule10nG[quad], M.rule10[quad].as<double>(), tDof, dAg, dDg,
                                 dVal, dR, kern, s);
    } else {
        set_error("phys %d not supported", phys);
        return SVFSI_ERR_NOTSUP;
    }
    return launch_asm_tet4(M, phys, to_asm(*p), make_tet4_rule(), tDof, dAg, dYg, dDg, dprop_el, dVal, dR, s);
}

int svfsi_assemble_traction_dev(void* mesh, int faceId, int dof, const double* h, double* dR, void* stream)
{
    SV_ARG(mesh && h && dR, "NULL argument");
    SV_TRY(require_device());
    Mesh& M = *static_cast<Mesh*>(mesh);
    SV_ARG(faceId >= 0 && faceId < (int)M.faceSets.size(), "bad faceId");
    SV_ARG(dof >= 3, "traction needs dof >= 3");
    return launch_face_traction(M, *M.faceSets[faceId], dof, h, make_tri3_rule(), dR,
                                static_cast<hipStream_t>(stream));
}

int svfsi_assemble_backflow_dev(void* mesh, int faceId, const svfsi_eq_params* p, int tDof, const double* dYg,
                                double* dVal, double* dR, void* stream)
{
    SV_ARG(mesh && p && dYg && dVal && dR, "NULL argument");
    SV_TRY(require_device());
    Mesh& M = *static_cast<Mesh*>(mesh);
    SV_ARG(faceId >= 0 && faceId < (int)M.faceSets.size(), "bad faceId");
    SV_ARG(tDof >= 4, "backflow needs tDof >= 4");
    return launch_face_backflow(M, *M.faceSets[faceId], to_asm(*p), tDof, dYg, make_tri3_rule(), dVal, dR,
                                static_cast<hipStream_t>(stream));
}

int svfsi_lelas_dkde_dot_dev(void* mesh, const svfsi_eq_params* p, const double* dlam, const double* du,
                             double* dg, void* stream)
{
    SV_ARG(mesh && p && dlam && du && dg, "NULL argument");
    SV_TRY(require_device());
    Mesh& M = *static_cast<Mesh*>(mesh);
    SV_ARG(M.eNoN == 4, "tet4 only");
    return launch_lelas_dkde(M, to_asm(*p), make_tet4_rule(), dlam, du, dg, static_cast<hipStream_t>(stream));
}

int svfsi_fluid_dmu_dot_dev(void* mesh, const svfsi_eq_params* p, int tDof, const double* dAg, const double* dYg,
                            const double* dmu_el, const double* dlam, const double* ddu, double* dg, void* stream)
{
    SV_ARG(mesh && p && dAg && dYg && dlam && ddu && dg, "NULL argument");
    SV_TRY(require_device());
    Mesh& M = *static_cast<Mesh*>(mesh);
    SV_ARG(M.eNoN == 4, "tet4 only");
    SV_ARG(tDof >= 4, "FLUID needs tDof >= 4");
    SV_ARG(p->viscType == 0 || p->viscType == SVFSI_VISC_CONST,
           "the viscosity sensitivity is defined for the constant (Newtonian) model only");
    return launch_fluid_dmu(M, to_asm(*p), make_tet4_rule(), tDof, dAg, dYg, dmu_el, dlam, ddu, dg,
                            static_cast<hipStream_t>(stream));
}

int svfsi_poly_factors(int d, int ld, const double* Hu, double* coef, int* kind, int maxFactors)
{
    SV_ARG(d >= 1 && ld >= d + 1 && Hu && coef && kind && maxFactors >= 1, "bad argument");
    std::vector<double> c;
    std::vector<int> k;
    const int nf = poly_factors(d, ld, Hu, c, k);
    if (nf < 0) {
        set_error("no usable harmonic Ritz values");
        return SVFSI_ERR_STATE;
    }
    SV_ARG(nf <= maxFactors, "%d factors exceed maxFactors %d", nf, maxFactors);
    std::copy(c.begin(), c.end(), coef);
    std::copy(k.begin(), k.end(), kind);
    return nf;
}

int svfsi_lhsa(int nNo, int nEl, int eNoN, const int* IEN, int* nnz, int* rowPtr, int* colPtr)
{
    SV_ARG(nNo > 0 && nEl > 0 && eNoN > 0 && IEN && nnz && rowPtr, "bad argument");
    // node -> element incidence
    std::vector<int> cnt(nNo + 1, 0);
    for (size_t q = 0; q < (size_t)nEl * eNoN; ++q) {
        const int a = IEN[q] - 1;
        SV_ARG(a >= 0 && a < nNo, "IEN entry %zu = %d out of range", q + 1, IEN[q]);
        cnt[a + 1]++;
    }
    for (int a = 0; a < nNo; ++a) cnt[a + 1] += cnt[a];
    std::vector<int> inc(cnt[nNo]);
    {
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (int e = 0; e < nEl; ++e)
            for (int a = 0; a < eNoN; ++a) inc[pos[IEN[(size_t)e * eNoN + a] - 1]++] = e;
    }
    // row r: sorted unique nodes of the elements incident to r (ADDCOL,
    // LHSA.f:205-241, gives the same sorted unique list)
    std::vector<int> rowLen(nNo);
    std::vector<int> cols;
    std::vector<int> scratch;
    size_t total = 0;
    const bool fill = colPtr != nullptr;
    for (int r = 0; r < nNo; ++r) {
        scratch.clear();
        for (int q = cnt[r]; q < cnt[r + 1]; ++q) {
            const int e = inc[q];
            for (int b = 0; b < eNoN; ++b) scratch.push_back(IEN[(size_t)e * eNoN + b]);
        }
        std::sort(scratch.begin(), scratch.end());
        scratch.erase(std::unique(scratch.begin(), scratch.end()), scratch.end());
        if (scratch.empty()) {
            set_error("Node %d is isolated (LHSA.f:177)", r + 1);
            return SVFSI_ERR_ARG;
        }
        rowLen[r] = (int)scratch.size();
        if (fill) {
            for (int c : scratch) colPtr[total++] = c;
        } else {
            total += scratch.size();
        }
    }
    SV_ARG(total < 2147483647u, "nnz overflows int32");
    rowPtr[0] = 1;
    for (int r = 0; r < nNo; ++r) rowPtr[r + 1] = rowPtr[r] + rowLen[r];
    *nnz = (int)total;
    return SVFSI_OK;
}

int svfsi_prof_enable(int on)
{
    prof().on = on != 0;
    return SVFSI_OK;
}

int svfsi_prof_get(int kernel, long long* count, double* total_ms)
{
    SV_ARG(kernel >= 0 && kernel < Prof::kSlots, "kernel must be 0..%d", Prof::kSlots - 1);
    prof_resolve();
    if (count) *count = prof().count[kernel];
    if (total_ms) *total_ms = prof().ms[kernel];
    return SVFSI_OK;
}

int svfsi_prof_get_bytes(int kernel, double* bytes)
{
    SV_ARG(kernel >= 0 && kernel < Prof::kSlots, "kernel must be 0..%d", Prof::kSlots - 1);
    prof_resolve();
    if (bytes) *bytes = prof().bytes[kernel];
    return SVFSI_OK;
}

int svfsi_prof_reset(void)
{
    prof_resolve();
    Prof& P = prof();
    for (int k = 0; k < Prof::kSlots; ++k) {
        P.count[k] = 0;
        P.ms[k] = 0.0;
        P.bytes[k] = 0.0;
    }
    return SVFSI_OK;
}

} // extern "C"
