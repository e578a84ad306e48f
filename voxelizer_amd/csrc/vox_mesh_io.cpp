// Mesh readers (OBJ, PLY, STL, legacy VTK) and node / voxel array writers
// (legacy VTK structured points, raw).  Host code only.
//
// The reference's tester loads models through VTK's readers
// (vtkOBJReader / vtkPLYReader / vtkSTLReader / vtkPolyDataReader, chosen by
// file extension, tests/src/main.cpp:359-404) and copies the polydata's
// points (as float) and polygons into flat arrays (loadModel,
// tests/src/main.cpp:911-955).  These readers restate the published file
// formats with the same result: float vertices in file order and triangles
// in file order, with two deliberate differences from that copy loop:
// polygons with more than three corners are split into a fan (0, k, k+1)
// (loadModel writes their extra ids into the next triangle's slots), and
// cells with fewer than three corners are skipped.  STL files are read as
// VTK's STL reader reads them by default: exactly coincident corners are
// merged into one vertex (numbered in order of first appearance) and facets
// that collapse under the merge are dropped.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/voxelizer_amd.h"

namespace {

struct Err {
    char *buf;
    size_t len;
    vox_status set(vox_status st, const std::string &m) const {
        if (buf && len) {
            const size_t n = std::min(len - 1, m.size());
            std::memcpy(buf, m.data(), n);
            buf[n] = '\0';
        }
        return st;
    }
};

struct MeshData {
    std::vector<float> v;
    std::vector<uint32_t> t;
};

bool read_file(const char *path, std::string &data) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    data = ss.str();
    return true;
}

std::string lower(std::string s) {
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

bool ends_with(const std::string &s, const char *suf) {
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// Appends polygon `ids` (already zero-based and checked) as a fan.
void add_polygon(MeshData &M, const uint32_t *ids, size_t n) {
    for (size_t k = 1; k + 1 < n; k++) {
        M.t.push_back(ids[0]);
        M.t.push_back(ids[k]);
        M.t.push_back(ids[k + 1]);
    }
}

// ---- tokenizer over a text buffer -------------------------------------------
struct Text {
    const std::string &s;
    size_t p = 0;
    explicit Text(const std::string &s_) : s(s_) {}
    bool eof() const { return p >= s.size(); }
    // next line without its end of line; false at the end
    bool line(std::string &out) {
        if (p >= s.size()) return false;
        size_t e = s.find('\n', p);
        if (e == std::string::npos) e = s.size();
        size_t end = e;
        if (end > p && s[end - 1] == '\r') end--;
        out.assign(s, p, end - p);
        p = e + 1;
        return true;
    }
    // next whitespace-separated token
    bool token(std::string &out) {
        while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
        if (p >= s.size()) return false;
        size_t b = p;
        while (p < s.size() && !std::isspace((unsigned char)s[p])) p++;
        out.assign(s, b, p - b);
        return true;
    }
};

bool parse_double(const std::string &tok, double &v) {
    if (tok.empty()) return false;
    char *end = nullptr;
    errno = 0;
    v = std::strtod(tok.c_str(), &end);
    return end == tok.c_str() + tok.size();
}
bool parse_long(const std::string &tok, long long &v) {
    if (tok.empty()) return false;
    char *end = nullptr;
    errno = 0;
    v = std::strtoll(tok.c_str(), &end, 10);
    return end == tok.c_str() + tok.size() && errno == 0;
}

// ---- OBJ ----------------------------------------------------------------------
// "v x y z [w]" vertices, "f i[/t][/n] ..." faces with 1-based or negative
// (relative) indices; every other statement is ignored.
vox_status load_obj(const std::string &data, MeshData &M, const Err &E) {
    Text T(data);
    std::string ln;
    std::vector<uint32_t> ids;
    size_t lineno = 0;
    while (T.line(ln)) {
        lineno++;
        std::istringstream ls(ln);
        std::string kw;
        if (!(ls >> kw) || kw[0] == '#') continue;
        if (kw == "v") {
            double c[3];
            std::string tok;
            for (int a = 0; a < 3; a++)
                if (!(ls >> tok) || !parse_double(tok, c[a]))
                    return E.set(VOX_ERR_INVALID, "OBJ line " + std::to_string(lineno) +
                                                      ": a vertex needs three numbers");
            for (int a = 0; a < 3; a++) M.v.push_back((float)c[a]);
        } else if (kw == "f") {
            ids.clear();
            std::string tok;
            const long long nv = (long long)(M.v.size() / 3);
            while (ls >> tok) {
                const std::string first = tok.substr(0, tok.find('/'));
                long long i;
                if (!parse_long(first, i) || i == 0)
                    return E.set(VOX_ERR_INVALID, "OBJ line " + std::to_string(lineno) +
                                                      ": bad face index '" + tok + "'");
                const long long z = i > 0 ? i - 1 : nv + i;
                if (z < 0 || z >= nv)
                    return E.set(VOX_ERR_INVALID, "OBJ line " + std::to_string(lineno) +
                                                      ": face index out of range");
                ids.push_back((uint32_t)z);
            }
            add_polygon(M, ids.data(), ids.size());
        }
    }
    return VOX_OK;
}

// ---- PLY ----------------------------------------------------------------------
enum PlyType { P_I8, P_U8, P_I16, P_U16, P_I32, P_U32, P_F32, P_F64, P_BAD };

PlyType ply_type(const std::string &t) {
    if (t == "char" || t == "int8") return P_I8;
    if (t == "uchar" || t == "uint8") return P_U8;
    if (t == "short" || t == "int16") return P_I16;
    if (t == "ushort" || t == "uint16") return P_U16;
    if (t == "int" || t == "int32") return P_I32;
    if (t == "uint" || t == "uint32") return P_U32;
    if (t == "float" || t == "float32") return P_F32;
    if (t == "double" || t == "float64") return P_F64;
    return P_BAD;
}
size_t ply_size(PlyType t) {
    static const size_t s[] = {1, 1, 2, 2, 4, 4, 4, 8, 0};
    return s[t];
}

struct PlyProp {
    std::string name;
    bool list = false;
    PlyType count = P_BAD, type = P_BAD;
};
struct PlyElem {
    std::string name;
    size_t n = 0;
    std::vector<PlyProp> props;
};

struct PlyBin {
    const std::string &s;
    size_t p;
    bool big;
    bool get(PlyType t, double &v) {
        const size_t n = ply_size(t);
        if (p + n > s.size()) return false;
        unsigned char b[8];
        std::memcpy(b, s.data() + p, n);
        p += n;
        if (big) std::reverse(b, b + n);
        switch (t) {
            case P_I8: { int8_t x; std::memcpy(&x, b, 1); v = x; break; }
            case P_U8: { uint8_t x; std::memcpy(&x, b, 1); v = x; break; }
            case P_I16: { int16_t x; std::memcpy(&x, b, 2); v = x; break; }
            case P_U16: { uint16_t x; std::memcpy(&x, b, 2); v = x; break; }
            case P_I32: { int32_t x; std::memcpy(&x, b, 4); v = x; break; }
            case P_U32: { uint32_t x; std::memcpy(&x, b, 4); v = x; break; }
            case P_F32: { float x; std::memcpy(&x, b, 4); v = x; break; }
            case P_F64: { double x; std::memcpy(&x, b, 8); v = x; break; }
            default: return false;
        }
        return true;
    }
};

vox_status load_ply(const std::string &data, MeshData &M, const Err &E) {
    Text T(data);
    std::string ln;
    if (!T.line(ln) || ln != "ply") return E.set(VOX_ERR_INVALID, "PLY: missing 'ply' magic");
    int format = -1;  // 0 ascii, 1 little endian, 2 big endian
    std::vector<PlyElem> elems;
    bool ended = false;
    while (T.line(ln)) {
        std::istringstream ls(ln);
        std::string kw;
        if (!(ls >> kw)) continue;
        if (kw == "format") {
            std::string f;
            ls >> f;
            format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1
                   : f == "binary_big_endian" ? 2 : -1;
            if (format < 0) return E.set(VOX_ERR_INVALID, "PLY: unknown format '" + f + "'");
        } else if (kw == "element") {
            PlyElem e;
            long long n;
            std::string ns;
            if (!(ls >> e.name >> ns) || !parse_long(ns, n) || n < 0)
                return E.set(VOX_ERR_INVALID, "PLY: bad element line");
            e.n = (size_t)n;
            elems.push_back(e);
        } else if (kw == "property") {
            if (elems.empty()) return E.set(VOX_ERR_INVALID, "PLY: property before element");
            PlyProp pr;
            std::string t;
            ls >> t;
            if (t == "list") {
                std::string ct, it;
                ls >> ct >> it >> pr.name;
                pr.list = true;
                pr.count = ply_type(ct);
                pr.type = ply_type(it);
                if (pr.count == P_BAD || pr.type == P_BAD)
                    return E.set(VOX_ERR_INVALID, "PLY: bad list property types");
            } else {
                pr.type = ply_type(t);
                ls >> pr.name;
                if (pr.type == P_BAD) return E.set(VOX_ERR_INVALID, "PLY: bad property type '" + t + "'");
            }
            elems.back().props.push_back(pr);
        } else if (kw == "end_header") {
            ended = true;
            break;
        }
    }
    if (!ended || format < 0) return E.set(VOX_ERR_INVALID, "PLY: incomplete header");
    PlyBin B{data, T.p, format == 2};
    std::vector<uint32_t> ids;
    size_t nverts = 0;
    for (const PlyElem &e : elems) {
        const bool is_v = e.name == "vertex", is_f = e.name == "face";
        int ix = -1, iy = -1, iz = -1, ifc = -1;
        for (size_t k = 0; k < e.props.size(); k++) {
            const std::string &n = e.props[k].name;
            if (is_v && n == "x") ix = (int)k;
            if (is_v && n == "y") iy = (int)k;
            if (is_v && n == "z") iz = (int)k;
            if (is_f && e.props[k].list && (n == "vertex_indices" || n == "vertex_index"))
                ifc = (int)k;
        }
        if (is_v && (ix < 0 || iy < 0 || iz < 0))
            return E.set(VOX_ERR_INVALID, "PLY: vertex element without x, y, z");
        if (is_v) nverts = e.n;
        for (size_t r = 0; r < e.n; r++) {
            double c[3] = {0, 0, 0};
            ids.clear();
            for (size_t k = 0; k < e.props.size(); k++) {
                const PlyProp &pr = e.props[k];
                size_t cnt = 1;
                if (pr.list) {
                    double cv;
                    if (format == 0) {
                        std::string tok;
                        if (!T.token(tok) || !parse_double(tok, cv))
                            return E.set(VOX_ERR_INVALID, "PLY: truncated data");
                    } else if (!B.get(pr.count, cv)) {
                        return E.set(VOX_ERR_INVALID, "PLY: truncated data");
                    }
                    if (cv < 0) return E.set(VOX_ERR_INVALID, "PLY: negative list length");
                    cnt = (size_t)cv;
                }
                for (size_t q = 0; q < cnt; q++) {
                    double x;
                    if (format == 0) {
                        std::string tok;
                        if (!T.token(tok) || !parse_double(tok, x))
                            return E.set(VOX_ERR_INVALID, "PLY: truncated data");
                    } else if (!B.get(pr.type, x)) {
                        return E.set(VOX_ERR_INVALID, "PLY: truncated data");
                    }
                    if ((int)k == ix) c[0] = x;
                    if ((int)k == iy) c[1] = x;
                    if ((int)k == iz) c[2] = x;
                    if ((int)k == ifc) {
                        if (x < 0 || x >= (double)nverts || x != std::floor(x))
                            return E.set(VOX_ERR_INVALID, "PLY: face index out of range");
                        ids.push_back((uint32_t)x);
                    }
                }
            }
            if (is_v)
                for (int a = 0; a < 3; a++) M.v.push_back((float)c[a]);
            if (is_f) add_polygon(M, ids.data(), ids.size());
        }
    }
    return VOX_OK;
}

// ---- STL ----------------------------------------------------------------------
struct VKey {
    uint32_t b[3];
    bool operator==(const VKey &o) const {
        return b[0] == o.b[0] && b[1] == o.b[1] && b[2] == o.b[2];
    }
};
struct VHash {
    size_t operator()(const VKey &k) const {
        uint64_t h = 1469598103934665603ull;
        for (int a = 0; a < 3; a++) h = (h ^ k.b[a]) * 1099511628211ull;
        return (size_t)h;
    }
};

struct StlMerge {
    MeshData &M;
    std::unordered_map<VKey, uint32_t, VHash> ids;
    explicit StlMerge(MeshData &m) : M(m) {}
    uint32_t add(const float c[3]) {
        VKey k;
        for (int a = 0; a < 3; a++) {
            // +0 and -0 are the same point
            const float x = c[a] == 0.0f ? 0.0f : c[a];
            std::memcpy(&k.b[a], &x, 4);
        }
        auto it = ids.find(k);
        if (it != ids.end()) return it->second;
        const uint32_t i = (uint32_t)(M.v.size() / 3);
        ids.emplace(k, i);
        for (int a = 0; a < 3; a++) M.v.push_back(c[a]);
        return i;
    }
    void facet(const float p[9]) {
        const uint32_t a = add(p), b = add(p + 3), c = add(p + 6);
        if (a != b && a != c && b != c) {
            M.t.push_back(a);
            M.t.push_back(b);
            M.t.push_back(c);
        }
    }
};

vox_status load_stl(const std::string &data, MeshData &M, const Err &E) {
    StlMerge S(M);
    if (data.size() >= 84) {
        uint32_t n;
        std::memcpy(&n, data.data() + 80, 4);
        if ((uint64_t)84 + 50ull * n == data.size()) {
            for (uint32_t f = 0; f < n; f++) {
                float p[9];
                std::memcpy(p, data.data() + 84 + 50ull * f + 12, 36);
                S.facet(p);
            }
            return VOX_OK;
        }
    }
    Text T(data);
    std::string tok;
    if (!T.token(tok) || lower(tok) != "solid")
        return E.set(VOX_ERR_INVALID, "STL: neither binary (size 84 + 50 n) nor ASCII ('solid')");
    float p[9];
    int k = 0;
    while (T.token(tok)) {
        const std::string t = lower(tok);
        if (t == "vertex") {
            if (k >= 9) return E.set(VOX_ERR_INVALID, "STL: more than three vertices in a facet");
            for (int a = 0; a < 3; a++) {
                double x;
                if (!T.token(tok) || !parse_double(tok, x))
                    return E.set(VOX_ERR_INVALID, "STL: bad vertex");
                p[k++] = (float)x;
            }
        } else if (t == "endloop") {
            if (k != 9) return E.set(VOX_ERR_INVALID, "STL: a facet needs three vertices");
            S.facet(p);
            k = 0;
        }
    }
    return VOX_OK;
}

// ---- legacy VTK polydata --------------------------------------------------------
size_t vtk_size(const std::string &t) {
    if (t == "float") return 4;
    if (t == "double") return 8;
    if (t == "int" || t == "unsigned_int") return 4;
    if (t == "long" || t == "unsigned_long" || t == "vtktypeint64" || t == "vtktypeuint64") return 8;
    if (t == "short" || t == "unsigned_short") return 2;
    if (t == "char" || t == "unsigned_char" || t == "bit") return 1;
    return 0;
}

struct VtkReader {
    const std::string &s;
    Text T;
    bool binary;
    explicit VtkReader(const std::string &d) : s(d), T(d), binary(false) {}
    bool number(const std::string &type, double &v) {
        if (!binary) {
            std::string tok;
            return T.token(tok) && parse_double(tok, v);
        }
        const size_t n = vtk_size(type);
        if (n == 0 || T.p + n > s.size()) return false;
        unsigned char b[8];
        std::memcpy(b, s.data() + T.p, n);
        T.p += n;
        std::reverse(b, b + n);  // legacy binary is big-endian
        if (type == "float") { float x; std::memcpy(&x, b, 4); v = x; }
        else if (type == "double") { double x; std::memcpy(&x, b, 8); v = x; }
        else if (type == "int") { int32_t x; std::memcpy(&x, b, 4); v = x; }
        else if (type == "unsigned_int") { uint32_t x; std::memcpy(&x, b, 4); v = x; }
        else if (type == "long" || type == "vtktypeint64") { int64_t x; std::memcpy(&x, b, 8); v = (double)x; }
        else if (type == "unsigned_long" || type == "vtktypeuint64") { uint64_t x; std::memcpy(&x, b, 8); v = (double)x; }
        else if (type == "short") { int16_t x; std::memcpy(&x, b, 2); v = x; }
        else if (type == "unsigned_short") { uint16_t x; std::memcpy(&x, b, 2); v = x; }
        else if (type == "char") { int8_t x; std::memcpy(&x, b, 1); v = x; }
        else { uint8_t x; std::memcpy(&x, b, 1); v = x; }
        return true;
    }
    // after binary data the next keyword starts on a fresh line
    void skip_eol() {
        if (binary)
            while (T.p < s.size() && (s[T.p] == '\n' || s[T.p] == '\r')) T.p++;
    }
};

vox_status load_vtk(const std::string &data, MeshData &M, const Err &E) {
    VtkReader R(data);
    std::string ln;
    if (!R.T.line(ln) || ln.rfind("# vtk DataFile", 0) != 0)
        return E.set(VOX_ERR_INVALID, "VTK: missing '# vtk DataFile' header");
    R.T.line(ln);  // title
    if (!R.T.line(ln)) return E.set(VOX_ERR_INVALID, "VTK: truncated header");
    const std::string fmt = lower(ln.substr(0, ln.find_first_of(" \t\r")));
    if (fmt != "ascii" && fmt != "binary") return E.set(VOX_ERR_INVALID, "VTK: format must be ASCII or BINARY");
    R.binary = fmt == "binary";
    std::string tok;
    if (!R.T.token(tok) || lower(tok) != "dataset" || !R.T.token(tok) || lower(tok) != "polydata")
        return E.set(VOX_ERR_INVALID, "VTK: only DATASET POLYDATA is read");
    size_t npts = 0;
    bool have_pts = false;
    std::vector<uint32_t> ids;
    while (R.T.token(tok)) {
        const std::string kw = lower(tok);
        if (kw == "points") {
            std::string ns, type;
            long long n;
            if (!R.T.token(ns) || !parse_long(ns, n) || n < 0 || !R.T.token(type))
                return E.set(VOX_ERR_INVALID, "VTK: bad POINTS line");
            R.T.line(ln);
            npts = (size_t)n;
            for (size_t i = 0; i < 3 * npts; i++) {
                double x;
                if (!R.number(type, x)) return E.set(VOX_ERR_INVALID, "VTK: truncated POINTS");
                M.v.push_back((float)x);
            }
            R.skip_eol();
            have_pts = true;
        } else if (kw == "polygons" || kw == "vertices" || kw == "lines" ||
                   kw == "triangle_strips") {
            std::string s1, s2;
            long long n, size;
            if (!R.T.token(s1) || !parse_long(s1, n) || !R.T.token(s2) || !parse_long(s2, size) ||
                n < 0 || size < 0)
                return E.set(VOX_ERR_INVALID, "VTK: bad cell section line");
            R.T.line(ln);
            const bool polys = kw == "polygons";
            // DataFile 5.1: OFFSETS <type> (n values) then CONNECTIVITY <type>
            // (size values); older files: n cells of "k id0 .. idk-1".
            const size_t save = R.T.p;
            std::string next;
            if (R.T.token(next) && lower(next) == "offsets") {
                std::string type;
                R.T.token(type);
                R.T.line(ln);
                std::vector<int64_t> off((size_t)n);
                for (auto &o : off) {
                    double x;
                    if (!R.number(type, x)) return E.set(VOX_ERR_INVALID, "VTK: truncated OFFSETS");
                    o = (int64_t)x;
                }
                R.skip_eol();
                if (!R.T.token(next) || lower(next) != "connectivity")
                    return E.set(VOX_ERR_INVALID, "VTK: OFFSETS without CONNECTIVITY");
                R.T.token(type);
                R.T.line(ln);
                std::vector<int64_t> con((size_t)size);
                for (auto &c : con) {
                    double x;
                    if (!R.number(type, x)) return E.set(VOX_ERR_INVALID, "VTK: truncated CONNECTIVITY");
                    c = (int64_t)x;
                }
                R.skip_eol();
                for (size_t c = 0; n > 0 && c + 1 < (size_t)n; c++) {
                    if (off[c] < 0 || off[c + 1] < off[c] || off[c + 1] > size)
                        return E.set(VOX_ERR_INVALID, "VTK: bad OFFSETS");
                    if (!polys) continue;
                    ids.clear();
                    for (int64_t q = off[c]; q < off[c + 1]; q++) {
                        if (con[q] < 0 || (size_t)con[q] >= npts)
                            return E.set(VOX_ERR_INVALID, "VTK: point id out of range");
                        ids.push_back((uint32_t)con[q]);
                    }
                    add_polygon(M, ids.data(), ids.size());
                }
                continue;
            }
            R.T.p = save;
            for (long long c = 0; c < n; c++) {
                double k;
                if (!R.number("int", k) || k < 0) return E.set(VOX_ERR_INVALID, "VTK: truncated cells");
                ids.clear();
                for (long long q = 0; q < (long long)k; q++) {
                    double id;
                    if (!R.number("int", id)) return E.set(VOX_ERR_INVALID, "VTK: truncated cells");
                    if (id < 0 || id >= (double)npts)
                        return E.set(VOX_ERR_INVALID, "VTK: point id out of range");
                    ids.push_back((uint32_t)id);
                }
                if (polys) add_polygon(M, ids.data(), ids.size());
            }
            R.skip_eol();
        } else if (kw == "point_data" || kw == "cell_data" || kw == "field") {
            break;  // attributes follow; the geometry is complete
        } else if (kw == "metadata") {
            // skip to the blank line that ends the block
            while (R.T.line(ln) && !ln.empty()) {
            }
        }
    }
    if (!have_pts) return E.set(VOX_ERR_INVALID, "VTK: no POINTS section");
    return VOX_OK;
}

void write_be(std::string &out, const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) out.push_back((char)b[n - 1 - i]);
}

}  // namespace

extern "C" {

vox_status vox_mesh_load(const char *path, vox_mesh *out, char *err, size_t err_len) {
    const Err E{err, err_len};
    if (err && err_len) err[0] = '\0';
    if (!path || !out) return E.set(VOX_ERR_INVALID, "null argument");
    out->vertices = nullptr;
    out->indices = nullptr;
    out->nv = out->nt = 0;
    const std::string p = lower(path);
    std::string data;
    if (!read_file(path, data)) return E.set(VOX_ERR_INVALID, std::string("cannot read ") + path);
    MeshData M;
    vox_status st;
    if (ends_with(p, ".obj")) st = load_obj(data, M, E);
    else if (ends_with(p, ".ply")) st = load_ply(data, M, E);
    else if (ends_with(p, ".stl")) st = load_stl(data, M, E);
    else if (ends_with(p, ".vtk")) st = load_vtk(data, M, E);
    else return E.set(VOX_ERR_UNSUPPORTED, "unknown mesh file extension (.obj .ply .stl .vtk)");
    if (st != VOX_OK) return st;
    if (M.v.size() / 3 > UINT32_MAX || M.t.size() / 3 > UINT32_MAX)
        return E.set(VOX_ERR_UNSUPPORTED, "mesh too large");
    out->nv = (uint32_t)(M.v.size() / 3);
    out->nt = (uint32_t)(M.t.size() / 3);
    out->vertices = (float *)std::malloc(std::max<size_t>(1, M.v.size()) * sizeof(float));
    out->indices = (uint32_t *)std::malloc(std::max<size_t>(1, M.t.size()) * sizeof(uint32_t));
    if (!out->vertices || !out->indices) {
        vox_mesh_free(out);
        return E.set(VOX_ERR_OOM, "out of host memory");
    }
    if (!M.v.empty()) std::memcpy(out->vertices, M.v.data(), M.v.size() * sizeof(float));
    if (!M.t.empty()) std::memcpy(out->indices, M.t.data(), M.t.size() * sizeof(uint32_t));
    return VOX_OK;
}

void vox_mesh_free(vox_mesh *m) {
    if (!m) return;
    std::free(m->vertices);
    std::free(m->indices);
    m->vertices = nullptr;
    m->indices = nullptr;
    m->nv = m->nt = 0;
}

vox_status vox_write_nodes_vtk(const char *path, const void *nodes, vox_node_kind kind,
                               const uint32_t dims[3], const double origin[3], double spacing,
                               int binary) {
    if (!path || !nodes || !dims) return VOX_ERR_INVALID;
    const size_t size = vox_node_size(kind);
    if (size == 0) return VOX_ERR_INVALID;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    const unsigned char *b = (const unsigned char *)nodes;
    std::string out;
    char hdr[512];
    std::snprintf(hdr, sizeof(hdr),
                  "# vtk DataFile Version 3.0\nvoxelizer_amd nodes (kind %d)\n%s\n"
                  "DATASET STRUCTURED_POINTS\nDIMENSIONS %u %u %u\nORIGIN %.17g %.17g %.17g\n"
                  "SPACING %.17g %.17g %.17g\nPOINT_DATA %zu\n",
                  (int)kind, binary ? "BINARY" : "ASCII", dims[0], dims[1], dims[2],
                  origin ? origin[0] : 0.0, origin ? origin[1] : 0.0, origin ? origin[2] : 0.0,
                  spacing, spacing, spacing, n);
    out += hdr;
    // bid and material of node i, decoded by the node layout
    auto bid = [&](size_t i) -> uint32_t {
        const unsigned char *q = b + i * size;
        switch (kind) {
            case VOX_NODE_SHORT: return q[0] & 31u;
            case VOX_NODE_LONG: case VOX_NODE_PARTIAL: return q[0];
            case VOX_NODE_SHORT_FCC: return (uint32_t)(q[0] | (q[1] << 8)) & 4095u;
            case VOX_NODE_LONG_FCC: return (uint32_t)(q[0] | (q[1] << 8));
        }
        return 0;
    };
    auto mat = [&](size_t i) -> uint32_t {
        const unsigned char *q = b + i * size;
        switch (kind) {
            case VOX_NODE_SHORT: return q[0] >> 5;
            case VOX_NODE_LONG: case VOX_NODE_PARTIAL: return q[1];
            case VOX_NODE_SHORT_FCC: return (uint32_t)(q[0] | (q[1] << 8)) >> 12;
            case VOX_NODE_LONG_FCC: return q[2];
        }
        return 0;
    };
    const bool fcc = kind == VOX_NODE_SHORT_FCC || kind == VOX_NODE_LONG_FCC;
    out += fcc ? "SCALARS bid unsigned_short 1\nLOOKUP_TABLE default\n"
               : "SCALARS bid unsigned_char 1\nLOOKUP_TABLE default\n";
    for (size_t i = 0; i < n; i++) {
        const uint32_t v = bid(i);
        if (binary) {
            if (fcc) {
                const uint16_t s = (uint16_t)v;
                write_be(out, &s, 2);
            } else {
                out.push_back((char)v);
            }
        } else {
            out += std::to_string(v);
            out.push_back('\n');
        }
    }
    if (binary) out.push_back('\n');
    out += "SCALARS material unsigned_char 1\nLOOKUP_TABLE default\n";
    for (size_t i = 0; i < n; i++) {
        const uint32_t v = mat(i);
        if (binary) {
            out.push_back((char)v);
        } else {
            out += std::to_string(v);
            out.push_back('\n');
        }
    }
    if (binary) out.push_back('\n');
    if (kind == VOX_NODE_PARTIAL) {
        out += "SCALARS ratio float 1\nLOOKUP_TABLE default\n";
        for (size_t i = 0; i < n; i++) {
            float r;
            std::memcpy(&r, b + i * size + 4, 4);
            if (binary) {
                write_be(out, &r, 4);
            } else {
                char t[32];
                std::snprintf(t, sizeof(t), "%.9g\n", r);
                out += t;
            }
        }
        if (binary) out.push_back('\n');
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) return VOX_ERR_INVALID;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return std::fclose(f) == 0 && ok ? VOX_OK : VOX_ERR_INVALID;
}

vox_status vox_write_bits_vtk(const char *path, const uint32_t *bits, const uint32_t dims[3],
                              const double origin[3], double spacing, int binary) {
    if (!path || !bits || !dims || dims[0] % 32 != 0) return VOX_ERR_INVALID;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    std::string out;
    char hdr[512];
    std::snprintf(hdr, sizeof(hdr),
                  "# vtk DataFile Version 3.0\nvoxelizer_amd solid voxels\n%s\n"
                  "DATASET STRUCTURED_POINTS\nDIMENSIONS %u %u %u\nORIGIN %.17g %.17g %.17g\n"
                  "SPACING %.17g %.17g %.17g\nPOINT_DATA %zu\n"
                  "SCALARS solid unsigned_char 1\nLOOKUP_TABLE default\n",
                  binary ? "BINARY" : "ASCII", dims[0], dims[1], dims[2],
                  origin ? origin[0] : 0.0, origin ? origin[1] : 0.0, origin ? origin[2] : 0.0,
                  spacing, spacing, spacing, n);
    out += hdr;
    for (size_t i = 0; i < n; i++) {
        // voxel x of a word sits at bit 31 - x % 32
        const uint32_t v = (bits[i / 32] >> (31u - (uint32_t)(i % 32))) & 1u;
        if (binary) {
            out.push_back((char)v);
        } else {
            out.push_back((char)('0' + v));
            out.push_back('\n');
        }
    }
    if (binary) out.push_back('\n');
    FILE *f = std::fopen(path, "wb");
    if (!f) return VOX_ERR_INVALID;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return std::fclose(f) == 0 && ok ? VOX_OK : VOX_ERR_INVALID;
}

}  // extern "C"
