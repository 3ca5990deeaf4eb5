// ply.cpp -- PLY mesh reader (host): ascii, binary_little_endian and
// binary_big_endian, any numeric property types, for LiAR's
// TriangleMesh::loadPly (src/liar/scenery/triangle_mesh.cpp:183-281).
//
// Reads the "vertex" element's x y z (and nx ny nz, u v when present) and the
// "face" element's vertex_indices (or vertex_index) lists; other elements and
// properties are skipped.  Faces become triangles as LiAR's loader makes them:
// its fan loop (i0, i_k, i_k+1) overwrites one triangle and stores it after
// the loop, so a polygon keeps only its last fan triangle (a quad 0 1 2 3
// gives 0 2 3).  LR_PLY_FAN_ALL keeps every fan triangle instead.
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/liar_amd.h"
#include "lr_util.h"

namespace {

enum class Type { I8, U8, I16, U16, I32, U32, F32, F64, BAD };

Type parse_type(const std::string& s) {
    if (s == "char" || s == "int8") return Type::I8;
    if (s == "uchar" || s == "uint8") return Type::U8;
    if (s == "short" || s == "int16") return Type::I16;
    if (s == "ushort" || s == "uint16") return Type::U16;
    if (s == "int" || s == "int32") return Type::I32;
    if (s == "uint" || s == "uint32") return Type::U32;
    if (s == "float" || s == "float32") return Type::F32;
    if (s == "double" || s == "float64") return Type::F64;
    return Type::BAD;
}

size_t type_size(Type t) {
    switch (t) {
    case Type::I8: case Type::U8: return 1;
    case Type::I16: case Type::U16: return 2;
    case Type::I32: case Type::U32: case Type::F32: return 4;
    case Type::F64: return 8;
    default: return 0;
    }
}

struct Property {
    std::string name;
    bool list = false;
    Type count = Type::BAD, item = Type::BAD;
};

struct Element {
    std::string name;
    uint64_t n = 0;
    std::vector<Property> props;
};

// Byte source over a whole file (read at once: PLY meshes of millions of
// triangles are tens of MB).
struct Reader {
    std::vector<unsigned char> buf;
    size_t pos = 0;
    bool swap = false;  // big-endian data on this little-endian host

    bool eof() const { return pos >= buf.size(); }

    bool line(std::string& out) {
        if (eof()) return false;
        size_t e = pos;
        while (e < buf.size() && buf[e] != '\n') ++e;
        out.assign(reinterpret_cast<const char*>(buf.data()) + pos, e - pos);
        if (!out.empty() && out.back() == '\r') out.pop_back();
        pos = e < buf.size() ? e + 1 : e;
        return true;
    }

    bool binary(Type t, double& v) {
        const size_t sz = type_size(t);
        if (pos + sz > buf.size()) return false;
        unsigned char b[8];
        memcpy(b, buf.data() + pos, sz);
        pos += sz;
        if (swap)
            for (size_t k = 0; k < sz / 2; ++k) std::swap(b[k], b[sz - 1 - k]);
        switch (t) {
        case Type::I8: { int8_t x; memcpy(&x, b, 1); v = x; break; }
        case Type::U8: { uint8_t x; memcpy(&x, b, 1); v = x; break; }
        case Type::I16: { int16_t x; memcpy(&x, b, 2); v = x; break; }
        case Type::U16: { uint16_t x; memcpy(&x, b, 2); v = x; break; }
        case Type::I32: { int32_t x; memcpy(&x, b, 4); v = x; break; }
        case Type::U32: { uint32_t x; memcpy(&x, b, 4); v = x; break; }
        case Type::F32: { float x; memcpy(&x, b, 4); v = x; break; }
        case Type::F64: { double x; memcpy(&x, b, 8); v = x; break; }
        default: return false;
        }
        return true;
    }

    // next whitespace-separated number (ascii body)
    bool ascii(double& v) {
        while (pos < buf.size() && (buf[pos] == ' ' || buf[pos] == '\t' || buf[pos] == '\r' || buf[pos] == '\n')) ++pos;
        if (eof()) return false;
        const char* s = reinterpret_cast<const char*>(buf.data()) + pos;
        size_t e = pos;
        while (e < buf.size() && !(buf[e] == ' ' || buf[e] == '\t' || buf[e] == '\r' || buf[e] == '\n')) ++e;
        std::string tok(s, e - pos);
        char* end = nullptr;
        errno = 0;
        v = strtod(tok.c_str(), &end);
        if (end == tok.c_str() || *end != '\0' || errno == ERANGE) return false;
        pos = e;
        return true;
    }
};

int bad(const std::string& path, const std::string& what) {
    return lr::fail(LR_E_INVALID, "lr_load_ply(" + path + "): " + what);
}

// Fewest body bytes one record of `e` can take: its fixed-size values and
// list counts in binary, one byte per number in ascii (each number is at
// least one character).  Counts in the header are checked against it before
// anything is sized by them, so a short file cannot ask for large buffers.
uint64_t min_record_bytes(const Element& e, bool ascii) {
    uint64_t b = 0;
    for (const Property& pr : e.props) b += ascii ? 1 : type_size(pr.list ? pr.count : pr.item);
    return b;
}

int load_ply(const char* path, uint32_t flags, lr_mesh_data* out) {
    const std::string p(path);
    Reader R;
    {
        FILE* f = fopen(path, "rb");
        if (!f) return bad(p, "could not open file");
        fseek(f, 0, SEEK_END);
        const long size = ftell(f);
        fseek(f, 0, SEEK_SET);
        if (size < 0) {
            fclose(f);
            return bad(p, "could not read file");
        }
        R.buf.resize((size_t)size);
        const size_t got = size ? fread(R.buf.data(), 1, (size_t)size, f) : 0;
        fclose(f);
        if (got != (size_t)size) return bad(p, "could not read file");
    }
    // header
    std::string ln;
    if (!R.line(ln) || ln != "ply") return bad(p, "not a PLY file");
    int fmt = -1;  // 0 ascii, 1 little, 2 big
    std::vector<Element> els;
    for (;;) {
        if (!R.line(ln)) return bad(p, "header without end_header");
        char w0[64] = {0}, w1[64] = {0}, w2[64] = {0}, w3[64] = {0}, w4[64] = {0};
        const int nw = sscanf(ln.c_str(), "%63s %63s %63s %63s %63s", w0, w1, w2, w3, w4);
        if (nw <= 0) continue;
        const std::string k(w0);
        if (k == "end_header") break;
        if (k == "comment" || k == "obj_info") continue;
        if (k == "format") {
            const std::string f(w1);
            fmt = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
            if (fmt < 0) return bad(p, "unknown format " + f);
        } else if (k == "element") {
            if (nw < 3) return bad(p, "bad element line");
            Element e;
            e.name = w1;
            e.n = strtoull(w2, nullptr, 10);
            els.push_back(e);
        } else if (k == "property") {
            if (els.empty()) return bad(p, "property before any element");
            Property pr;
            if (std::string(w1) == "list") {
                if (nw < 5) return bad(p, "bad list property");
                pr.list = true;
                pr.count = parse_type(w2);
                pr.item = parse_type(w3);
                pr.name = w4;
                if (pr.count == Type::BAD || pr.item == Type::BAD) return bad(p, "bad list property types");
            } else {
                if (nw < 3) return bad(p, "bad property line");
                pr.item = parse_type(w1);
                pr.name = w2;
                if (pr.item == Type::BAD) return bad(p, std::string("unknown property type ") + w1);
            }
            els.back().props.push_back(pr);
        } else {
            return bad(p, "unknown header line: " + ln);
        }
    }
    if (fmt < 0) return bad(p, "no format line");
    R.swap = fmt == 2;
    auto num = [&](Type t, double& v) { return fmt == 0 ? R.ascii(v) : R.binary(t, v); };

    std::vector<float> verts, normals, uvs;
    std::vector<uint32_t> tris;
    bool have_vertex = false;
    uint64_t nv = 0;
    const bool fan_all = (flags & LR_PLY_FAN_ALL) != 0;
    for (const Element& e : els) {
        const bool is_v = e.name == "vertex", is_f = e.name == "face";
        int ix = -1, iy = -1, iz = -1, inx = -1, iny = -1, inz = -1, iu = -1, iv = -1, ifc = -1;
        for (size_t k = 0; k < e.props.size(); ++k) {
            const std::string& n = e.props[k].name;
            if (is_v && !e.props[k].list) {
                if (n == "x") ix = (int)k;
                if (n == "y") iy = (int)k;
                if (n == "z") iz = (int)k;
                if (n == "nx") inx = (int)k;
                if (n == "ny") iny = (int)k;
                if (n == "nz") inz = (int)k;
                if (n == "u") iu = (int)k;
                if (n == "v") iv = (int)k;
            }
            if (is_f && e.props[k].list && (n == "vertex_indices" || (n == "vertex_index" && ifc < 0))) ifc = (int)k;
        }
        if (is_v) {
            if (ix < 0 || iy < 0 || iz < 0) return bad(p, "vertex element without x, y, z");
            const uint64_t left = R.buf.size() - std::min(R.pos, R.buf.size());
            const uint64_t rec = min_record_bytes(e, fmt == 0);
            if (rec && e.n > left / rec) return bad(p, "vertex element has more records than the file holds");
            if (e.n >= (1ull << 32)) return bad(p, "too many vertices");
            have_vertex = true;
            nv = e.n;
            verts.resize(3 * e.n);
            if (inx >= 0 && iny >= 0 && inz >= 0) normals.resize(3 * e.n);
            if (iu >= 0 && iv >= 0) uvs.resize(2 * e.n);
        }
        if (is_f && ifc < 0) return bad(p, "face element without vertex_indices");
        {
            const uint64_t left = R.buf.size() - std::min(R.pos, R.buf.size());
            const uint64_t rec = min_record_bytes(e, fmt == 0);
            if (rec && e.n > left / rec) return bad(p, "element " + e.name + " has more records than the file holds");
        }
        std::vector<double> row(e.props.size());
        std::vector<uint64_t> idx;
        for (uint64_t r = 0; r < e.n; ++r) {
            for (size_t k = 0; k < e.props.size(); ++k) {
                const Property& pr = e.props[k];
                if (!pr.list) {
                    if (!num(pr.item, row[k])) return bad(p, "truncated or malformed " + e.name + " data");
                    continue;
                }
                double c;
                if (!num(pr.count, c) || c < 0 || c != std::floor(c)) return bad(p, "bad list count in " + e.name);
                const uint64_t cnt = (uint64_t)c;
                {
                    const uint64_t left = R.buf.size() - std::min(R.pos, R.buf.size());
                    const uint64_t item = fmt == 0 ? 1 : type_size(pr.item);
                    if (cnt > left / item) return bad(p, "list in " + e.name + " is longer than the file");
                }
                if ((int)k == ifc) idx.resize(cnt);
                for (uint64_t q = 0; q < cnt; ++q) {
                    double x;
                    if (!num(pr.item, x)) return bad(p, "truncated list in " + e.name);
                    if ((int)k == ifc) {
                        if (x < 0 || x != std::floor(x)) return bad(p, "negative or fractional vertex index");
                        idx[q] = (uint64_t)x;
                    }
                }
            }
            if (is_v) {
                verts[3 * r + 0] = (float)row[ix];
                verts[3 * r + 1] = (float)row[iy];
                verts[3 * r + 2] = (float)row[iz];
                if (!normals.empty()) {
                    normals[3 * r + 0] = (float)row[inx];
                    normals[3 * r + 1] = (float)row[iny];
                    normals[3 * r + 2] = (float)row[inz];
                }
                if (!uvs.empty()) {
                    uvs[2 * r + 0] = (float)row[iu];
                    uvs[2 * r + 1] = (float)row[iv];
                }
            } else if (is_f) {
                if (idx.size() < 3) return bad(p, "faces must have at least 3 vertices");
                for (uint64_t q : idx)
                    if (!have_vertex || q >= nv) return bad(p, "vertex index out of range");
                const size_t last = idx.size() - 2;
                for (size_t k2 = fan_all ? 1 : last; k2 <= last; ++k2) {
                    tris.push_back((uint32_t)idx[0]);
                    tris.push_back((uint32_t)idx[k2]);
                    tris.push_back((uint32_t)idx[k2 + 1]);
                }
            }
        }
    }
    if (!have_vertex) return bad(p, "no vertex element");
    out->n_verts = (uint32_t)nv;
    out->n_tris = (uint32_t)(tris.size() / 3);
    auto copy_out = [](const auto& v, auto** dst) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        *dst = nullptr;
        if (v.empty()) return true;
        *dst = static_cast<T*>(malloc(sizeof(T) * v.size()));
        if (!*dst) return false;
        memcpy(*dst, v.data(), sizeof(T) * v.size());
        return true;
    };
    if (!copy_out(verts, &out->verts) || !copy_out(normals, &out->normals) || !copy_out(uvs, &out->uvs) ||
        !copy_out(tris, &out->tris)) {
        lr_mesh_data_free(out);
        return lr::fail(LR_E_NOMEM, "lr_load_ply: out of host memory");
    }
    return LR_OK;
}

}  // namespace

extern "C" {

int lr_load_ply(const char* path, uint32_t flags, lr_mesh_data* out) {
    if (!path || !out) return lr::fail(LR_E_INVALID, "lr_load_ply: NULL argument");
    memset(out, 0, sizeof(*out));
    try {
        return load_ply(path, flags, out);
    } catch (const std::bad_alloc&) {
        lr_mesh_data_free(out);
        return lr::fail(LR_E_NOMEM, std::string("lr_load_ply(") + path + "): out of host memory");
    } catch (const std::exception& ex) {
        lr_mesh_data_free(out);
        return lr::fail(LR_E_INVALID, std::string("lr_load_ply(") + path + "): " + ex.what());
    }
}

void lr_mesh_data_free(lr_mesh_data* m) {
    if (!m) return;
    free(m->verts);
    free(m->normals);
    free(m->uvs);
    free(m->tris);
    memset(m, 0, sizeof(*m));
}

}  // extern "C"
