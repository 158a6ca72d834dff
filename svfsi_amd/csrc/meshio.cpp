// meshio.cpp -- the input side of the path: svFSI's coordinate/connectivity
// mesh files (READCCNE, svFSI/LOADMSH.f:125-239), their face files
// ("Connectivity file (ebc) path", LOADMSH.f:196-216) and the element
// orientation check CHECKIEN (svFSI/READMSH.f:1010-1185) for tet4/tet10.
// Host code: the files are read once, before anything reaches the device.
#include "internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <utility>
#include <string>
#include <vector>

namespace svfsi {

namespace {

// The numbers of a text line, as list-directed input reads them
// (separators: blanks, tabs, commas).
bool parse_line(const std::string& line, std::vector<double>& out)
{
    out.clear();
    const char* p = line.c_str();
    for (;;) {
        while (*p == ' ' || *p == '\t' || *p == ',' || *p == '\r' || *p == '\n') ++p;
        if (!*p) break;
        char* end = nullptr;
        const double v = std::strtod(p, &end);
        if (end == p) return false;
        out.push_back(v);
        p = end;
    }
    return true;
}

// Every non-blank line of a file as numbers; width = numbers on the first.
int read_table(const char* path, std::vector<std::vector<double>>& rows)
{
    rows.clear();
    FILE* f = std::fopen(path, "r");
    if (!f) {
        set_error("cannot open %s", path);
        return SVFSI_ERR_ARG;
    }
    std::string line;
    std::vector<double> v;
    char buf[4096];
    int lineNo = 0;
    bool ok = true;
    while (ok && std::fgets(buf, sizeof(buf), f)) {
        line = buf;
        while (!line.empty() && line.back() != '\n' && std::fgets(buf, sizeof(buf), f)) line += buf;
        ++lineNo;
        if (!parse_line(line, v)) {
            set_error("%s:%d: not a list of numbers", path, lineNo);
            ok = false;
            break;
        }
        if (v.empty()) continue;
        if (!rows.empty() && v.size() != rows[0].size()) {
            set_error("%s:%d: %zu numbers, the first line has %zu", path, lineNo, v.size(), rows[0].size());
            ok = false;
            break;
        }
        rows.push_back(v);
    }
    std::fclose(f);
    if (!ok) return SVFSI_ERR_ARG;
    if (rows.empty()) {
        set_error("%s holds no data", path);
        return SVFSI_ERR_ARG;
    }
    return SVFSI_OK;
}

int to_int(double v, int* out)
{
    if (!(std::fabs(v - std::nearbyint(v)) == 0.0) || std::fabs(v) > 2147483647.0) return SVFSI_ERR_ARG;
    *out = (int)v;
    return SVFSI_OK;
}

int sgn(double v) { return v > 0.0 ? 1 : (v < 0.0 ? -1 : 0); }

} // namespace

} // namespace svfsi

using namespace svfsi;

extern "C" {

int svfsi_read_ccne_sizes(const char* connPath, const char* coordPath, int* nEl, int* eNoN, int* nNo, int* nsd)
{
    SV_ARG(connPath && coordPath && nEl && eNoN && nNo && nsd, "NULL argument");
    std::vector<std::vector<double>> c, x;
    SV_TRY(read_table(connPath, c));
    SV_TRY(read_table(coordPath, x));
    SV_ARG(c[0].size() >= 2 && x[0].size() >= 2, "connectivity needs an index and nodes, coordinates an index "
                                                  "and components");
    *nEl = (int)c.size();
    *eNoN = (int)c[0].size() - 1;
    *nNo = (int)x.size();
    *nsd = (int)x[0].size() - 1;
    return SVFSI_OK;
}

int svfsi_read_ccne(const char* connPath, const char* coordPath, int nEl, int eNoN, int nNo, int nsd, int* ien,
                    double* x)
{
    SV_ARG(connPath && coordPath && ien && x, "NULL argument");
    std::vector<std::vector<double>> c, xs;
    SV_TRY(read_table(connPath, c));
    SV_TRY(read_table(coordPath, xs));
    SV_ARG((int)c.size() == nEl && (int)c[0].size() == eNoN + 1, "connectivity file does not match nEl/eNoN");
    SV_ARG((int)xs.size() == nNo && (int)xs[0].size() == nsd + 1, "coordinates file does not match nNo/nsd");
    for (int e = 0; e < nEl; ++e)
        for (int a = 0; a < eNoN; ++a)
            if (to_int(c[e][a + 1], &ien[(size_t)e * eNoN + a]) != SVFSI_OK) {
                set_error("element %d: node %d is not an integer", e + 1, a + 1);
                return SVFSI_ERR_ARG;
            }
    for (int i = 0; i < nNo; ++i)
        for (int d = 0; d < nsd; ++d) x[(size_t)i * nsd + d] = xs[i][d + 1];
    return SVFSI_OK;
}

int svfsi_read_ebc_sizes(const char* path, int* nFe, int* feNoN)
{
    SV_ARG(path && nFe && feNoN, "NULL argument");
    std::vector<std::vector<double>> r;
    SV_TRY(read_table(path, r));
    SV_ARG(r[0].size() >= 3, "a face line holds the element, an index and the face nodes");
    *nFe = (int)r.size();
    *feNoN = (int)r[0].size() - 2;
    return SVFSI_OK;
}

int svfsi_read_ebc(const char* path, int nFe, int feNoN, int* gE, int* fien)
{
    SV_ARG(path && gE && fien, "NULL argument");
    std::vector<std::vector<double>> r;
    SV_TRY(read_table(path, r));
    SV_ARG((int)r.size() == nFe && (int)r[0].size() == feNoN + 2, "face file does not match nFe/feNoN");
    for (int e = 0; e < nFe; ++e) {
        bool ok = to_int(r[e][0], &gE[e]) == SVFSI_OK;
        for (int a = 0; a < feNoN && ok; ++a) ok = to_int(r[e][a + 2], &fien[(size_t)e * feNoN + a]) == SVFSI_OK;
        if (!ok) {
            set_error("face element %d: not integers", e + 1);
            return SVFSI_ERR_ARG;
        }
    }
    return SVFSI_OK;
}

int svfsi_checkien(int eNoN, int nEl, int nNo, int* ien, const double* x, int* nReordered)
{
    SV_ARG(ien && x, "NULL argument");
    SV_ARG(eNoN == 4 || eNoN == 10, "CHECKIEN here covers tet4 and tet10 (eNoN 4 or 10)");
    std::vector<char> used(nNo, 0);
    for (int e = 0; e < nEl; ++e)
        for (int a = 0; a < eNoN; ++a) {
            const int Ac = ien[(size_t)e * eNoN + a];
            if (Ac > nNo || Ac <= 0) {
                set_error("Element %d contains an out of bound node", e + 1);
                return SVFSI_ERR_ARG;
            }
            used[Ac - 1] = 1;
        }
    for (int i = 0; i < nNo; ++i)
        if (!used[i]) {
            set_error("Node %d is isolated from the mesh", i + 1);
            return SVFSI_ERR_ARG;
        }
    int swapped = 0;
    for (int e = 0; e < nEl; ++e) {
        int* n = ien + (size_t)e * eNoN;
        const double* p[4];
        for (int a = 0; a < 4; ++a) p[a] = x + (size_t)(n[a] - 1) * 3;
        double v1[3], v2[3], v3[3];
        for (int d = 0; d < 3; ++d) {
            v1[d] = p[1][d] - p[0][d];
            v2[d] = p[2][d] - p[1][d];
            v3[d] = p[3][d] - p[2][d];
        }
        const double v4[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                              v1[0] * v2[1] - v1[1] * v2[0]};
        const int s = sgn(v3[0] * v4[0] + v3[1] * v4[1] + v3[2] * v4[2]);
        if (s == 0) {
            set_error("Element %d is distorted", e + 1);
            return SVFSI_ERR_ARG;
        }
        if (s == 1) {
            std::swap(n[0], n[1]);
            if (eNoN == 10) {
                std::swap(n[5], n[6]);
                std::swap(n[7], n[8]);
            }
            ++swapped;
        }
    }
    if (nReordered) *nReordered = swapped;
    return SVFSI_OK;
}

} // extern "C"
