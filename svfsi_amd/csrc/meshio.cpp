// meshio.cpp -- the input side of the path: svFSI's coordinate/connectivity
// mesh files (READCCNE, svFSI/LOADMSH.f:125-239), their face files
// ("Connectivity file (ebc) path", LOADMSH.f:196-216) and the element
// orientation check CHECKIEN (svFSI/READMSH.f:1010-1185) for tet4/tet10.
// Host code: the files are read once, before anything reaches the device.
#include "internal.hpp"

#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
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

// ---- VTK XML (.vtu UnstructuredGrid, .vtp PolyData), the files svFSI's
// vtkXMLParser reads for meshes and faces (svFSI/VTKXML.f) ----

struct Tag {
    size_t begin = std::string::npos, end = 0; // tag text [begin, end)
    size_t body = 0, bodyEnd = 0;              // element content
    std::string attrs;
};

std::string attr(const std::string& attrs, const char* name)
{
    const std::string key = std::string(" ") + name + "=\"";
    size_t p = (" " + attrs).find(key);
    if (p == std::string::npos) return "";
    p += key.size() - 1; // index in attrs (the added space shifts by one)
    const size_t q = attrs.find('"', p);
    return q == std::string::npos ? "" : attrs.substr(p, q - p);
}

// next <name ...> at or after `from`, with its content up to </name>
Tag find_tag(const std::string& s, const char* name, size_t from, size_t limit)
{
    Tag t;
    const std::string open = std::string("<") + name;
    for (size_t p = s.find(open, from); p != std::string::npos && p < limit; p = s.find(open, p + 1)) {
        const char c = s[p + open.size()];
        if (c != ' ' && c != '>' && c != '/' && c != '\n' && c != '\t' && c != '\r') continue;
        const size_t gt = s.find('>', p);
        if (gt == std::string::npos) break;
        t.begin = p;
        t.attrs = s.substr(p + open.size(), gt - p - open.size());
        for (char& ch : t.attrs)
            if (ch == '\n' || ch == '\t' || ch == '\r') ch = ' ';
        t.body = gt + 1;
        if (s[gt - 1] == '/') {
            t.bodyEnd = t.body;
            t.end = gt + 1;
        } else {
            const size_t close = s.find(std::string("</") + name, gt);
            t.bodyEnd = close == std::string::npos ? s.size() : close;
            t.end = t.bodyEnd;
        }
        return t;
    }
    return t;
}

int b64val(char c)
{
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+') return 62;
    if (c == '/') return 63;
    return -1;
}

// base64 in groups of four characters, each group decoded on its own, so
// separately encoded streams that were concatenated (VTK's compressed
// header and blocks) decode too; whitespace is skipped.
std::vector<uint8_t> b64decode(const char* p, const char* e)
{
    std::vector<uint8_t> out;
    int q[4], k = 0;
    for (; p < e; ++p) {
        if (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t') continue;
        q[k++] = *p == '=' ? -2 : b64val(*p);
        if (k < 4) continue;
        k = 0;
        if (q[0] < 0 || q[1] < 0) break;
        out.push_back((uint8_t)((q[0] << 2) | (q[1] >> 4)));
        if (q[2] >= 0) out.push_back((uint8_t)(((q[1] & 15) << 4) | (q[2] >> 2)));
        if (q[2] >= 0 && q[3] >= 0) out.push_back((uint8_t)(((q[2] & 3) << 6) | q[3]));
    }
    return out;
}

size_t type_size(const std::string& t)
{
    if (t == "Int8" || t == "UInt8") return 1;
    if (t == "Int16" || t == "UInt16") return 2;
    if (t == "Int32" || t == "UInt32" || t == "Float32") return 4;
    if (t == "Int64" || t == "UInt64" || t == "Float64") return 8;
    return 0;
}

double typed_at(const uint8_t* b, const std::string& t)
{
    if (t == "Float64") { double v; std::memcpy(&v, b, 8); return v; }
    if (t == "Float32") { float v; std::memcpy(&v, b, 4); return v; }
    if (t == "Int64") { int64_t v; std::memcpy(&v, b, 8); return (double)v; }
    if (t == "UInt64") { uint64_t v; std::memcpy(&v, b, 8); return (double)v; }
    if (t == "Int32") { int32_t v; std::memcpy(&v, b, 4); return v; }
    if (t == "UInt32") { uint32_t v; std::memcpy(&v, b, 4); return v; }
    if (t == "Int16") { int16_t v; std::memcpy(&v, b, 2); return v; }
    if (t == "UInt16") { uint16_t v; std::memcpy(&v, b, 2); return v; }
    if (t == "Int8") return (double)(int8_t)b[0];
    return (double)b[0];
}

struct VtkFile {
    std::string s, path;
    size_t hdr = 4;        // header integer size (UInt32 or UInt64)
    bool zlib = false;
    size_t app = std::string::npos; // first byte after the '_' of AppendedData
    bool appB64 = false;
};

uint64_t read_uint(const uint8_t* b, size_t n)
{
    if (n == 8) { uint64_t v; std::memcpy(&v, b, 8); return v; }
    uint32_t v; std::memcpy(&v, b, 4); return v;
}

// Largest decoded array accepted (bytes).  The sizes in a binary block's
// header come from the file; bounding them keeps a corrupt or crafted header
// from asking for an allocation the process cannot make.
constexpr uint64_t kMaxArrayBytes = uint64_t(1) << 36;

// payload bytes of a binary block (header + data, optionally zlib blocks).
// Every size read from the block is checked against the bytes available
// before it is used in an offset, in a form that cannot overflow.
int unpack_block(const VtkFile& f, const uint8_t* b, size_t avail, std::vector<uint8_t>& out, size_t* used)
{
    const size_t h = f.hdr;
    if (!f.zlib) {
        if (avail < h) return SVFSI_ERR_ARG;
        const uint64_t n = read_uint(b, h);
        if (n > avail - h) return SVFSI_ERR_ARG;
        out.assign(b + h, b + h + n);
        if (used) *used = h + n;
        return SVFSI_OK;
    }
    if (avail < 3 * h) return SVFSI_ERR_ARG;
    const uint64_t nb = read_uint(b, h), bs = read_uint(b + h, h), last = read_uint(b + 2 * h, h);
    // the nb compressed sizes follow the three header integers
    if (nb > avail / h - 3) return SVFSI_ERR_ARG;
    if (bs > kMaxArrayBytes || last > bs || (nb && nb - 1 > kMaxArrayBytes / (bs ? bs : 1))) return SVFSI_ERR_ARG;
    size_t off = (3 + nb) * h;
    out.clear();
    for (uint64_t k = 0; k < nb; ++k) {
        const uint64_t cs = read_uint(b + (3 + k) * h, h);
        const uint64_t us = (k + 1 == nb && last) ? last : bs;
        if (cs > avail - off) return SVFSI_ERR_ARG;
        if (us > kMaxArrayBytes - out.size()) return SVFSI_ERR_ARG;
        const size_t o = out.size();
        out.resize(o + us);
        uLongf dl = (uLongf)us;
        if (uncompress(out.data() + o, &dl, b + off, (uLong)cs) != Z_OK || dl != us) return SVFSI_ERR_ARG;
        off += cs;
    }
    if (used) *used = off;
    return SVFSI_OK;
}

// the values of one DataArray as doubles
int array_values(const VtkFile& f, const Tag& t, std::vector<double>& v)
{
    const std::string type = attr(t.attrs, "type"), fmt = attr(t.attrs, "format");
    const size_t ts = type_size(type);
    if (!ts) {
        set_error("%s: DataArray type '%s' not supported", f.path.c_str(), type.c_str());
        return SVFSI_ERR_ARG;
    }
    v.clear();
    std::vector<uint8_t> raw;
    if (fmt == "ascii") {
        std::vector<double> tmp;
        if (!parse_line(f.s.substr(t.body, t.bodyEnd - t.body), tmp)) {
            set_error("%s: bad ascii DataArray", f.path.c_str());
            return SVFSI_ERR_ARG;
        }
        v = tmp;
        return SVFSI_OK;
    } else if (fmt == "binary") {
        // one base64 stream, or (compressed) the header's and the blocks'
        // streams one after the other: b64decode takes both
        const std::vector<uint8_t> enc = b64decode(f.s.data() + t.body, f.s.data() + t.bodyEnd);
        if (unpack_block(f, enc.data(), enc.size(), raw, nullptr) != SVFSI_OK) {
            set_error("%s: bad binary DataArray", f.path.c_str());
            return SVFSI_ERR_ARG;
        }
    } else if (fmt == "appended") {
        if (f.app == std::string::npos) {
            set_error("%s: appended DataArray without AppendedData", f.path.c_str());
            return SVFSI_ERR_ARG;
        }
        const size_t off = (size_t)std::strtoull(attr(t.attrs, "offset").c_str(), nullptr, 10);
        if (off > f.s.size() - f.app) {
            set_error("%s: appended DataArray offset beyond the end of the file", f.path.c_str());
            return SVFSI_ERR_ARG;
        }
        if (f.appB64) {
            size_t e = f.s.find('<', f.app + off);
            if (e == std::string::npos) e = f.s.size();
            const std::vector<uint8_t> enc = b64decode(f.s.data() + f.app + off, f.s.data() + e);
            if (unpack_block(f, enc.data(), enc.size(), raw, nullptr) != SVFSI_OK) {
                set_error("%s: bad appended DataArray", f.path.c_str());
                return SVFSI_ERR_ARG;
            }
        } else {
            const uint8_t* b = reinterpret_cast<const uint8_t*>(f.s.data()) + f.app + off;
            if (unpack_block(f, b, f.s.size() - f.app - off, raw, nullptr) != SVFSI_OK) {
                set_error("%s: bad appended DataArray", f.path.c_str());
                return SVFSI_ERR_ARG;
            }
        }
    } else {
        set_error("%s: DataArray format '%s' not supported", f.path.c_str(), fmt.c_str());
        return SVFSI_ERR_ARG;
    }
    v.resize(raw.size() / ts);
    for (size_t k = 0; k < v.size(); ++k) v[k] = typed_at(raw.data() + k * ts, type);
    return SVFSI_OK;
}

int vtk_open(const char* path, VtkFile& f)
{
    FILE* fp = std::fopen(path, "rb");
    if (!fp) {
        set_error("cannot open %s", path);
        return SVFSI_ERR_ARG;
    }
    std::fseek(fp, 0, SEEK_END);
    const long n = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    f.s.resize(n > 0 ? (size_t)n : 0);
    const size_t got = n > 0 ? std::fread(&f.s[0], 1, (size_t)n, fp) : 0;
    std::fclose(fp);
    f.path = path;
    if (got != f.s.size()) {
        set_error("cannot read %s", path);
        return SVFSI_ERR_ARG;
    }
    const Tag vf = find_tag(f.s, "VTKFile", 0, f.s.size());
    if (vf.begin == std::string::npos) {
        set_error("%s: not a VTK XML file", path);
        return SVFSI_ERR_ARG;
    }
    const std::string bo = attr(vf.attrs, "byte_order");
    if (!bo.empty() && bo != "LittleEndian") {
        set_error("%s: big-endian data not supported", path);
        return SVFSI_ERR_ARG;
    }
    f.hdr = attr(vf.attrs, "header_type") == "UInt64" ? 8 : 4;
    const std::string comp = attr(vf.attrs, "compressor");
    if (!comp.empty() && comp != "vtkZLibDataCompressor") {
        set_error("%s: compressor %s not supported", path, comp.c_str());
        return SVFSI_ERR_ARG;
    }
    f.zlib = !comp.empty();
    const Tag ad = find_tag(f.s, "AppendedData", 0, f.s.size());
    if (ad.begin != std::string::npos) {
        f.appB64 = attr(ad.attrs, "encoding") == "base64";
        const size_t u = f.s.find('_', ad.body);
        if (u != std::string::npos) f.app = u + 1;
    }
    return SVFSI_OK;
}

// DataArray named `name` inside the first <section> (empty name: the first)
int section_array(const VtkFile& f, const char* section, const char* name, std::vector<double>& v, bool* found)
{
    if (found) *found = false;
    const Tag sec = find_tag(f.s, section, 0, f.app == std::string::npos ? f.s.size() : f.app);
    if (sec.begin == std::string::npos) return SVFSI_OK;
    for (size_t p = sec.body;;) {
        const Tag da = find_tag(f.s, "DataArray", p, sec.bodyEnd);
        if (da.begin == std::string::npos) break;
        if (!*name || attr(da.attrs, "Name") == name) {
            if (found) *found = true;
            return array_values(f, da, v);
        }
        p = da.end;
    }
    return SVFSI_OK;
}

struct VtkMesh {
    std::vector<double> x, conn, offs, gnid, geid;
    bool hasN = false, hasE = false;
    int npc = 0, nCells = 0;
};

int vtk_mesh(const char* path, VtkMesh& m)
{
    VtkFile f;
    SV_TRY(vtk_open(path, f));
    bool ok = false;
    SV_TRY(section_array(f, "Points", "", m.x, &ok));
    if (!ok || m.x.size() % 3) {
        set_error("%s: no 3-component Points array", path);
        return SVFSI_ERR_ARG;
    }
    const bool poly = find_tag(f.s, "PolyData", 0, f.s.size()).begin != std::string::npos;
    const char* cells = poly ? "Polys" : "Cells";
    SV_TRY(section_array(f, cells, "connectivity", m.conn, &ok));
    bool ok2 = false;
    SV_TRY(section_array(f, cells, "offsets", m.offs, &ok2));
    if (!ok || !ok2 || m.offs.empty()) {
        set_error("%s: no %s connectivity/offsets", path, cells);
        return SVFSI_ERR_ARG;
    }
    // offsets come from the file: check them before any conversion to int
    if (m.offs.size() > (size_t)INT32_MAX || m.x.size() / 3 > (size_t)INT32_MAX) {
        set_error("%s: too many cells or points", path);
        return SVFSI_ERR_ARG;
    }
    for (const double o : m.offs)
        if (!(o >= 1.0 && o <= (double)m.conn.size())) {
            set_error("%s: cell offset out of range", path);
            return SVFSI_ERR_ARG;
        }
    m.nCells = (int)m.offs.size();
    m.npc = (int)m.offs[0];
    for (int c = 0; c < m.nCells; ++c)
        if ((int)m.offs[c] != (c + 1) * m.npc) {
            set_error("%s: cells of different sizes", path);
            return SVFSI_ERR_ARG;
        }
    if (m.conn.size() != (size_t)m.nCells * (size_t)m.npc) {
        set_error("%s: connectivity size does not match the offsets", path);
        return SVFSI_ERR_ARG;
    }
    SV_TRY(section_array(f, "PointData", "GlobalNodeID", m.gnid, &m.hasN));
    SV_TRY(section_array(f, "CellData", "GlobalElementID", m.geid, &m.hasE));
    return SVFSI_OK;
}

} // namespace

} // namespace svfsi

using namespace svfsi;

namespace {


int svfsi_read_ccne_sizes_body(const char* connPath, const char* coordPath, int* nEl, int* eNoN, int* nNo, int* nsd)
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

int svfsi_read_ccne_body(const char* connPath, const char* coordPath, int nEl, int eNoN, int nNo, int nsd, int* ien,
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

int svfsi_read_ebc_sizes_body(const char* path, int* nFe, int* feNoN)
{
    SV_ARG(path && nFe && feNoN, "NULL argument");
    std::vector<std::vector<double>> r;
    SV_TRY(read_table(path, r));
    SV_ARG(r[0].size() >= 3, "a face line holds the element, an index and the face nodes");
    *nFe = (int)r.size();
    *feNoN = (int)r[0].size() - 2;
    return SVFSI_OK;
}

int svfsi_read_ebc_body(const char* path, int nFe, int feNoN, int* gE, int* fien)
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

int svfsi_read_vtk_sizes_body(const char* path, int* nPoints, int* nCells, int* nodesPerCell, int* hasNodeIds,
                         int* hasElementIds)
{
    SV_ARG(path && nPoints && nCells && nodesPerCell, "NULL argument");
    VtkMesh m;
    SV_TRY(vtk_mesh(path, m));
    *nPoints = (int)(m.x.size() / 3);
    *nCells = m.nCells;
    *nodesPerCell = m.npc;
    if (hasNodeIds) *hasNodeIds = m.hasN;
    if (hasElementIds) *hasElementIds = m.hasE;
    return SVFSI_OK;
}

int svfsi_read_vtk_body(const char* path, int nPoints, int nCells, int nodesPerCell, double* x, int* conn,
                   int* nodeIds, int* elementIds)
{
    SV_ARG(path && x && conn, "NULL argument");
    VtkMesh m;
    SV_TRY(vtk_mesh(path, m));
    SV_ARG((int)(m.x.size() / 3) == nPoints && m.nCells == nCells && m.npc == nodesPerCell,
           "sizes do not match the file");
    for (size_t k = 0; k < m.x.size(); ++k) x[k] = m.x[k];
    for (size_t k = 0; k < m.conn.size(); ++k) {
        SV_ARG(m.conn[k] >= 0 && m.conn[k] < nPoints, "cell node out of range");
        conn[k] = (int)m.conn[k];
    }
    if (nodeIds) {
        SV_ARG(m.hasN && (int)m.gnid.size() == nPoints, "no GlobalNodeID point array");
        for (int k = 0; k < nPoints; ++k)
            SV_ARG(to_int(m.gnid[k], &nodeIds[k]) == SVFSI_OK, "GlobalNodeID %d is not an integer", k + 1);
    }
    if (elementIds) {
        SV_ARG(m.hasE && (int)m.geid.size() == nCells, "no GlobalElementID cell array");
        for (int k = 0; k < nCells; ++k)
            SV_ARG(to_int(m.geid[k], &elementIds[k]) == SVFSI_OK, "GlobalElementID %d is not an integer", k + 1);
    }
    return SVFSI_OK;
}

int svfsi_checkien_body(int eNoN, int nEl, int nNo, int* ien, const double* x, int* nReordered)
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


} // namespace

extern "C" {

int svfsi_read_ccne_sizes(const char* connPath, const char* coordPath, int* nEl, int* eNoN, int* nNo, int* nsd)
{
    return guarded([&] { return svfsi_read_ccne_sizes_body(connPath, coordPath, nEl, eNoN, nNo, nsd); });
}

int svfsi_read_ccne(const char* connPath, const char* coordPath, int nEl, int eNoN, int nNo, int nsd, int* ien,
                    double* x)
{
    return guarded([&] { return svfsi_read_ccne_body(connPath, coordPath, nEl, eNoN, nNo, nsd, ien, x); });
}

int svfsi_read_ebc_sizes(const char* path, int* nFe, int* feNoN)
{
    return guarded([&] { return svfsi_read_ebc_sizes_body(path, nFe, feNoN); });
}

int svfsi_read_ebc(const char* path, int nFe, int feNoN, int* gE, int* fien)
{
    return guarded([&] { return svfsi_read_ebc_body(path, nFe, feNoN, gE, fien); });
}

int svfsi_read_vtk_sizes(const char* path, int* nPoints, int* nCells, int* nodesPerCell, int* hasNodeIds,
                         int* hasElementIds)
{
    return guarded([&] { return svfsi_read_vtk_sizes_body(path, nPoints, nCells, nodesPerCell, hasNodeIds, hasElementIds); });
}

int svfsi_read_vtk(const char* path, int nPoints, int nCells, int nodesPerCell, double* x, int* conn,
                   int* nodeIds, int* elementIds)
{
    return guarded([&] { return svfsi_read_vtk_body(path, nPoints, nCells, nodesPerCell, x, conn, nodeIds, elementIds); });
}

int svfsi_checkien(int eNoN, int nEl, int nNo, int* ien, const double* x, int* nReordered)
{
    return guarded([&] { return svfsi_checkien_body(eNoN, nEl, nNo, ien, x, nReordered); });
}

} // extern "C"
