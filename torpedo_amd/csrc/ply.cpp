// 3DGS PLY reader: GaussianPoint::fromModel (torpedo/volumetric/src/GaussianGeometry.cpp:59-127).
// The reference parses the file with the third-party miniply; this is a reader of its own for
// the PLY header grammar (ascii, binary_little_endian, binary_big_endian; scalar and list
// properties of every PLY type) that extracts the vertex element's 3DGS columns and converts
// them as the reference does:
//   position  x, y, z
//   opacity   1 / (1 + exp(-opacity))                               (:112)
//   rotation  (rot_1, rot_2, rot_3, rot_0) normalised (w first in the file)   (:113-114)
//   scale     exp(scale_0..2), modifier w = 1                        (:115)
//   sh        f_dc_0..2 then the properties following f_dc_2 in file order, as many as there
//             are f_rest_* properties (:52-58, :85-88); the rest of the 48 stay zero.
#include "tpd_ctx.hpp"

#include <torpedo/math.hpp>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <stdexcept>
#include <sstream>
#include <string>
#include <vector>

namespace tpd {
namespace {

enum class PlyType { I8, U8, I16, U16, I32, U32, F32, F64, Bad };

PlyType ply_type(const std::string& t) {
    if (t == "char" || t == "int8") return PlyType::I8;
    if (t == "uchar" || t == "uint8") return PlyType::U8;
    if (t == "short" || t == "int16") return PlyType::I16;
    if (t == "ushort" || t == "uint16") return PlyType::U16;
    if (t == "int" || t == "int32") return PlyType::I32;
    if (t == "uint" || t == "uint32") return PlyType::U32;
    if (t == "float" || t == "float32") return PlyType::F32;
    if (t == "double" || t == "float64") return PlyType::F64;
    return PlyType::Bad;
}

size_t type_size(PlyType t) {
    switch (t) {
        case PlyType::I8: case PlyType::U8: return 1;
        case PlyType::I16: case PlyType::U16: return 2;
        case PlyType::I32: case PlyType::U32: case PlyType::F32: return 4;
        case PlyType::F64: return 8;
        default: return 0;
    }
}

struct Property {
    std::string name;
    PlyType type = PlyType::Bad;
    bool is_list = false;
    PlyType count_type = PlyType::Bad;
};

struct Element {
    std::string name;
    uint64_t count = 0;
    std::vector<Property> props;
};

enum class Format { Ascii, BinLE, BinBE };

struct Reader {
    const std::vector<char>& buf;
    size_t pos;
    Format fmt;

    bool read_raw(size_t n, unsigned char* out) {
        if (pos + n > buf.size()) return false;
        std::memcpy(out, buf.data() + pos, n);
        pos += n;
        const bool host_le = [] { const uint16_t x = 1; return *reinterpret_cast<const unsigned char*>(&x) == 1; }();
        if ((fmt == Format::BinBE) == host_le)
            for (size_t i = 0; i < n / 2; ++i) std::swap(out[i], out[n - 1 - i]);
        return true;
    }

    bool next_token(std::string& tok) {
        while (pos < buf.size() && std::isspace(static_cast<unsigned char>(buf[pos]))) ++pos;
        if (pos >= buf.size()) return false;
        const size_t start = pos;
        while (pos < buf.size() && !std::isspace(static_cast<unsigned char>(buf[pos]))) ++pos;
        tok.assign(buf.data() + start, pos - start);
        return true;
    }

    // Reads one value of type t as double (exact for every PLY type).
    bool value(PlyType t, double& v) {
        if (fmt == Format::Ascii) {
            std::string tok;
            if (!next_token(tok)) return false;
            char* end = nullptr;
            v = std::strtod(tok.c_str(), &end);
            return end && *end == '\0';
        }
        unsigned char b[8];
        const size_t n = type_size(t);
        if (n == 0 || !read_raw(n, b)) return false;
        switch (t) {
            case PlyType::I8: { int8_t x; std::memcpy(&x, b, 1); v = x; break; }
            case PlyType::U8: { uint8_t x; std::memcpy(&x, b, 1); v = x; break; }
            case PlyType::I16: { int16_t x; std::memcpy(&x, b, 2); v = x; break; }
            case PlyType::U16: { uint16_t x; std::memcpy(&x, b, 2); v = x; break; }
            case PlyType::I32: { int32_t x; std::memcpy(&x, b, 4); v = x; break; }
            case PlyType::U32: { uint32_t x; std::memcpy(&x, b, 4); v = x; break; }
            case PlyType::F32: { float x; std::memcpy(&x, b, 4); v = x; break; }
            case PlyType::F64: { double x; std::memcpy(&x, b, 8); v = x; break; }
            default: return false;
        }
        return true;
    }
};

int load(const char* path, std::vector<float>& out, uint32_t& n_out) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f.is_open()) return fail(TPD_ERR_INVALID, std::string("Failed to open file: ") + path);
    const std::streamsize size = f.tellg();
    if (size < 0) return fail(TPD_ERR_INVALID, std::string("Failed to read file: ") + path);
    f.seekg(0);
    std::vector<char> buf((size_t)size);
    if (size > 0 && !f.read(buf.data(), size)) return fail(TPD_ERR_INVALID, std::string("Failed to read file: ") + path);

    // header
    size_t pos = 0;
    auto next_line = [&](std::string& line) -> bool {
        if (pos >= buf.size()) return false;
        size_t e = pos;
        while (e < buf.size() && buf[e] != '\n') ++e;
        line.assign(buf.data() + pos, e - pos);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        pos = e < buf.size() ? e + 1 : e;
        return true;
    };
    std::string line;
    if (!next_line(line) || line != "ply") return fail(TPD_ERR_INVALID, std::string("not a PLY file: ") + path);
    Format fmt = Format::Ascii;
    bool have_format = false, end = false;
    std::vector<Element> elements;
    while (next_line(line)) {
        std::istringstream ls(line);
        std::string kw;
        ls >> kw;
        if (kw == "format") {
            std::string name, ver;
            ls >> name >> ver;
            if (name == "ascii") fmt = Format::Ascii;
            else if (name == "binary_little_endian") fmt = Format::BinLE;
            else if (name == "binary_big_endian") fmt = Format::BinBE;
            else return fail(TPD_ERR_INVALID, "PLY: unknown format " + name);
            have_format = true;
        } else if (kw == "element") {
            Element el;
            ls >> el.name >> el.count;
            if (ls.fail()) return fail(TPD_ERR_INVALID, "PLY: bad element line: " + line);
            elements.push_back(el);
        } else if (kw == "property") {
            if (elements.empty()) return fail(TPD_ERR_INVALID, "PLY: property before element");
            Property p;
            std::string t;
            ls >> t;
            if (t == "list") {
                std::string ct, it;
                ls >> ct >> it >> p.name;
                p.is_list = true;
                p.count_type = ply_type(ct);
                p.type = ply_type(it);
                if (p.count_type == PlyType::Bad || p.type == PlyType::Bad || p.count_type == PlyType::F32 ||
                    p.count_type == PlyType::F64)
                    return fail(TPD_ERR_INVALID, "PLY: bad list property: " + line);
            } else {
                ls >> p.name;
                p.type = ply_type(t);
                if (p.type == PlyType::Bad) return fail(TPD_ERR_INVALID, "PLY: bad property type: " + line);
            }
            elements.back().props.push_back(p);
        } else if (kw == "end_header") {
            end = true;
            break;
        }  // comment, obj_info: ignored
    }
    if (!have_format || !end) return fail(TPD_ERR_INVALID, std::string("PLY: incomplete header: ") + path);

    int vi = -1;
    for (size_t i = 0; i < elements.size(); ++i)
        if (elements[i].name == "vertex") { vi = (int)i; break; }
    if (vi < 0) return fail(TPD_ERR_INVALID, std::string("Could NOT find vertices: ") + path);
    const Element& vert = elements[vi];
    if (vert.count > 0xFFFFFFFFull) return fail(TPD_ERR_CAPACITY, "PLY: too many vertices");

    auto find = [&](const char* name) -> int {
        for (size_t i = 0; i < vert.props.size(); ++i)
            if (!vert.props[i].is_list && vert.props[i].name == name) return (int)i;
        return -1;
    };
    const char* need[] = {"x", "y", "z", "opacity", "rot_0", "rot_1", "rot_2", "rot_3", "scale_0", "scale_1",
                          "scale_2", "f_dc_0", "f_dc_1", "f_dc_2"};
    int col[14];
    for (int i = 0; i < 14; ++i) {
        col[i] = find(need[i]);
        if (col[i] < 0) return fail(TPD_ERR_INVALID, std::string("PLY: vertex property missing: ") + need[i]);
    }
    uint32_t rest = 0;
    for (const auto& p : vert.props)
        if (p.name.rfind("f_rest_", 0) == 0) ++rest;
    const uint32_t features = std::min<uint32_t>(rest + 3, 48);
    // features follow f_dc_0 contiguously in property order (GaussianGeometry.cpp:85-88)
    if ((size_t)col[11] + features > vert.props.size())
        return fail(TPD_ERR_INVALID, "PLY: SH properties are not contiguous after f_dc_0");
    for (uint32_t k = 0; k < features; ++k)
        if (vert.props[col[11] + k].is_list) return fail(TPD_ERR_INVALID, "PLY: list property among SH features");

    Reader rd{buf, pos, fmt};
    std::vector<double> row;
    auto skip_or_read = [&](const Element& el, bool keep, std::vector<double>* values) -> bool {
        if (keep) values->assign(el.props.size(), 0.0);
        for (size_t i = 0; i < el.props.size(); ++i) {
            const Property& p = el.props[i];
            if (p.is_list) {
                double cnt;
                if (!rd.value(p.count_type, cnt) || cnt < 0) return false;
                for (uint64_t k = 0; k < (uint64_t)cnt; ++k) {
                    double dummy;
                    if (!rd.value(p.type, dummy)) return false;
                }
            } else {
                double v;
                if (!rd.value(p.type, v)) return false;
                if (keep) (*values)[i] = v;
            }
        }
        return true;
    };
    // The counts come from the file: before sizing anything by them, check that the data after the
    // header can hold that many vertex rows (each value takes at least its binary size, or one
    // character and a separator in ASCII).
    uint64_t min_row = 0;
    for (const auto& p : vert.props)
        min_row += fmt == Format::Ascii ? 2 : type_size(p.is_list ? p.count_type : p.type);
    if (vert.count > 0 && (min_row == 0 || vert.count > (buf.size() - pos + 1) / min_row))
        return fail(TPD_ERR_INVALID, std::string("PLY: truncated data (the vertex count exceeds what the file holds): ") +
                                         path);
    const uint32_t n = (uint32_t)vert.count;
    out.assign((size_t)n * 60, 0.0f);
    for (size_t ei = 0; ei < elements.size(); ++ei) {
        const Element& el = elements[ei];
        const bool is_vert = (int)ei == vi;
        for (uint64_t r = 0; r < el.count; ++r) {
            if (!skip_or_read(el, is_vert, &row)) return fail(TPD_ERR_INVALID, std::string("PLY: truncated data: ") + path);
            if (!is_vert) continue;
            float* g = &out[(size_t)r * 60];
            g[0] = (float)row[col[0]];
            g[1] = (float)row[col[1]];
            g[2] = (float)row[col[2]];
            const float op = (float)row[col[3]];
            g[3] = 1.0f / (1.0f + std::exp(-op));
            const vec4 q{(float)row[col[5]], (float)row[col[6]], (float)row[col[7]], (float)row[col[4]]};
            const float inv = 1.0f / std::sqrt(math::dot(q, q));
            g[4] = q.x * inv;
            g[5] = q.y * inv;
            g[6] = q.z * inv;
            g[7] = q.w * inv;
            g[8] = std::exp((float)row[col[8]]);
            g[9] = std::exp((float)row[col[9]]);
            g[10] = std::exp((float)row[col[10]]);
            g[11] = 1.0f;
            for (uint32_t k = 0; k < features; ++k) g[12 + k] = (float)row[col[11] + k];
        }
        if (is_vert) break;  // elements after the vertices are not needed
    }
    n_out = n;
    return TPD_OK;
}

}  // namespace
}  // namespace tpd

using namespace tpd;

extern "C" {

int tpd_load_ply(const char* path, float** records, uint32_t* n) {
    if (!path || !records || !n) return fail(TPD_ERR_INVALID, "tpd_load_ply: bad argument");
    *records = nullptr;
    *n = 0;
    std::vector<float> out;
    uint32_t count = 0;
    int rc;
    try {
        rc = load(path, out, count);
    } catch (const std::bad_alloc&) {
        return fail(TPD_ERR_CAPACITY, std::string("tpd_load_ply: out of host memory reading ") + path);
    } catch (const std::exception& ex) {
        return fail(TPD_ERR_INVALID, std::string("tpd_load_ply: ") + ex.what());
    }
    if (rc != TPD_OK) return rc;
    float* p = static_cast<float*>(std::malloc(std::max<size_t>(out.size(), 1) * sizeof(float)));
    if (!p) return fail(TPD_ERR_CAPACITY, "tpd_load_ply: out of host memory");
    if (!out.empty()) std::memcpy(p, out.data(), out.size() * sizeof(float));
    *records = p;
    *n = count;
    return TPD_OK;
}

void tpd_free_host(void* p) { std::free(p); }

}  // extern "C"
