// Asset loaders of the reference's scene pipeline (host C++):
//   - Radiance RGBE (.hdr) environment maps, HDRLoader (src/loader/HDRLoader.cpp:5-153),
//     and a writer for tools and tests;
//   - Wavefront OBJ meshes with groups and JSON material libraries, OBJLoader
//     (src/loader/OBJLoader.cpp:3-105) + MaterialLoader (src/loader/MaterialLoader.cpp:9-36),
//     centred and scaled as Object::postProcess (src/core/Object.cpp:40-87);
//   - JSON scene configurations, Config (src/util/Config.cpp:3-22) + SceneLoader
//     (src/loader/SceneLoader.cpp:8-44) + Transform (src/geometry/Transform.cpp:13-34):
//     mesh objects with scale / translate / orientation, analytic planes and spheres,
//     the material table and the orbit camera.
// JSON objects iterate in key order, as the reference's std::map-backed JSON and
// MaterialPool do, so material indices match the reference's.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "tt_host.h"

namespace tt {
namespace {

// ------------------------------------------------------------------ JSON
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::map<std::string, JVal> obj;

    const JVal *get(const std::string &k) const {
        if (kind != OBJ) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

struct JParser {
    const std::string &s;
    size_t i = 0;
    std::string err;
    explicit JParser(const std::string &text) : s(text) {}

    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool fail(const char *m) {
        if (err.empty()) err = std::string(m) + " at offset " + std::to_string(i);
        return false;
    }
    bool lit(const char *w) {
        size_t n = std::strlen(w);
        if (s.compare(i, n, w) != 0) return fail("bad literal");
        i += n;
        return true;
    }
    bool string(std::string &out) {
        if (i >= s.size() || s[i] != '"') return fail("expected string");
        ++i;
        out.clear();
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) return fail("bad escape");
                char e = s[i++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (i + 4 > s.size()) return fail("bad \\u escape");
                        unsigned cp = (unsigned)std::stoul(s.substr(i, 4), nullptr, 16);
                        i += 4;
                        if (cp < 0x80) out += (char)cp;
                        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 63)); }
                        else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 63)); out += (char)(0x80 | (cp & 63)); }
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) return fail("unterminated string");
        ++i;
        return true;
    }
    bool value(JVal &v, int depth) {
        if (depth > 64) return fail("nesting too deep");
        ws();
        if (i >= s.size()) return fail("unexpected end");
        char c = s[i];
        if (c == '{') {
            v.kind = JVal::OBJ;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') { ++i; return true; }
            while (true) {
                ws();
                std::string k;
                if (!string(k)) return false;
                ws();
                if (i >= s.size() || s[i] != ':') return fail("expected ':'");
                ++i;
                JVal child;
                if (!value(child, depth + 1)) return false;
                v.obj[k] = std::move(child);
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == '}') { ++i; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JVal::ARR;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') { ++i; return true; }
            while (true) {
                JVal child;
                if (!value(child, depth + 1)) return false;
                v.arr.push_back(std::move(child));
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == ']') { ++i; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JVal::STR;
            return string(v.str);
        }
        if (c == 't') { v.kind = JVal::BOOL; v.b = true; return lit("true"); }
        if (c == 'f') { v.kind = JVal::BOOL; v.b = false; return lit("false"); }
        if (c == 'n') { v.kind = JVal::NUL; return lit("null"); }
        const char *start = s.c_str() + i;
        char *end = nullptr;
        v.num = std::strtod(start, &end);
        if (end == start) return fail("bad value");
        v.kind = JVal::NUM;
        i += (size_t)(end - start);
        return true;
    }
};

bool read_file(const std::string &path, std::string &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

tt_status parse_json_file(const std::string &path, JVal &v) {
    std::string text;
    if (!read_file(path, text)) return fail(TT_ERR_IO, "cannot read " + path);
    JParser p(text);
    if (!p.value(v, 0)) return fail(TT_ERR_IO, path + ": JSON " + p.err);
    return TT_OK;
}

bool jvec3(const JVal *v, float out[3]) {
    if (!v || v->kind != JVal::ARR || v->arr.size() < 3) return false;
    for (int k = 0; k < 3; ++k) {
        if (v->arr[k].kind != JVal::NUM) return false;
        out[k] = (float)v->arr[k].num;
    }
    return true;
}

// ------------------------------------------------------------------ materials
// One JSON material (Material.cpp:9-18, Coat/Diff/Metal/Spec/Refr.cpp): the
// type decides which colour fields are read; emission is optional.
tt_status json_material(const std::string &name, const JVal &j, tt_material &m) {
    m = tt_material{};
    m.ior = 1.4f;          // Refr: the reference's fixed index (RefrImpl.cuh)
    m.roughness = 0.3f;    // Metal in path mode (the reference's Phong lobe has no roughness)
    for (int k = 0; k < 3; ++k) m.specular[k] = 1.0f;
    const JVal *t = j.get("type");
    if (!t || t->kind != JVal::STR) return fail(TT_ERR_INVALID, "material " + name + ": no type");
    const JVal *e = j.get("emission");
    if (e && !jvec3(e, m.emission)) return fail(TT_ERR_INVALID, "material " + name + ": bad emission");
    bool ok = true;
    if (t->str == "Diff") {
        m.type = TT_MAT_DIFFUSE;
        ok = jvec3(j.get("diffuseColor"), m.albedo);
    } else if (t->str == "Metal") {
        m.type = TT_MAT_CONDUCTOR;
        ok = jvec3(j.get("color"), m.albedo);
    } else if (t->str == "Spec") {
        m.type = TT_MAT_SPECULAR;
        ok = jvec3(j.get("color"), m.albedo);
    } else if (t->str == "Refr") {
        m.type = TT_MAT_DIELECTRIC;
        m.albedo[0] = m.albedo[1] = m.albedo[2] = 1.0f;
    } else if (t->str == "Coat") {
        m.type = TT_MAT_COAT;
        ok = jvec3(j.get("diffuseColor"), m.albedo) && jvec3(j.get("specularColor"), m.specular);
    } else {
        return fail(TT_ERR_INVALID, "material " + name + ": unknown type " + t->str);
    }
    if (!ok) return fail(TT_ERR_INVALID, "material " + name + ": missing colour");
    return TT_OK;
}

// A material library: name -> material, in key order (MaterialPool's std::map).
struct Pool {
    std::vector<std::string> names;
    std::vector<tt_material> mats;
    int index(const std::string &n) const {
        auto it = std::lower_bound(names.begin(), names.end(), n);
        return (it != names.end() && *it == n) ? (int)(it - names.begin()) : -1;
    }
};

tt_status load_pool(const std::string &path, Pool &pool) {
    JVal root;
    tt_status st = parse_json_file(path, root);
    if (st != TT_OK) return st;
    if (root.kind != JVal::OBJ) return fail(TT_ERR_INVALID, path + ": material library must be a JSON object");
    for (const auto &kv : root.obj) {
        tt_material m;
        st = json_material(kv.first, kv.second, m);
        if (st != TT_OK) return st;
        pool.names.push_back(kv.first);
        pool.mats.push_back(m);
    }
    return TT_OK;
}

std::string dir_of(const std::string &p) {
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

// ------------------------------------------------------------------ OBJ
struct ObjGroup {
    std::string name, material;
    std::vector<int32_t> tv, tn;  // 3 per triangle, 0-based
};
struct Obj {
    std::vector<V3> verts, normals;
    std::vector<ObjGroup> groups;
    Pool pool;
};

// Whitespace-separated float triples of a "v" / "vn" line (the reference keeps
// reading triples until the line ends).
bool read_triples(std::istringstream &ls, std::vector<V3> &out) {
    float x, y, z;
    while (ls >> x >> y >> z) out.push_back(V3{x, y, z});
    return true;
}

tt_status load_obj(const std::string &path, Obj &o) {
    std::ifstream f(path);
    if (!f.good()) return fail(TT_ERR_IO, "cannot read " + path);
    std::string line;
    int64_t lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        std::istringstream ls(line);
        std::string key;
        ls >> key;
        if (key == "g") {
            ObjGroup g;
            ls >> g.name;
            o.groups.push_back(g);
        } else if (key == "mtllib") {
            std::string name;
            ls >> name;
            o.pool = Pool();
            tt_status st = load_pool(dir_of(path) + name, o.pool);
            if (st != TT_OK) return st;
        } else if (key == "usemtl") {
            std::string name;
            ls >> name;
            if (o.pool.index(name) < 0) return fail(TT_ERR_INVALID, path + ": material " + name + " not found");
            if (o.groups.empty()) o.groups.push_back(ObjGroup());
            o.groups.back().material = name;
        } else if (key == "v") {
            read_triples(ls, o.verts);
        } else if (key == "vn") {
            read_triples(ls, o.normals);
        } else if (key == "f") {
            // v, v/t, v//n, v/t/n; 1-based; a fan over the polygon
            std::vector<int32_t> vi, ni;
            std::string tok;
            int32_t v = 1, n = 1;
            while (ls >> tok) {
                size_t s1 = tok.find('/');
                v = (int32_t)std::strtol(tok.c_str(), nullptr, 10);
                if (s1 != std::string::npos) {
                    size_t s2 = tok.find('/', s1 + 1);
                    if (s2 != std::string::npos && s2 + 1 < tok.size())
                        n = (int32_t)std::strtol(tok.c_str() + s2 + 1, nullptr, 10);
                }
                if (v <= 0) return fail(TT_ERR_INVALID, path + ": line " + std::to_string(lineno) +
                                                            ": only positive vertex indices are supported");
                vi.push_back(v - 1);
                ni.push_back(n - 1);
            }
            if (vi.size() < 3) return fail(TT_ERR_INVALID, path + ": line " + std::to_string(lineno) + ": face with < 3 vertices");
            if (o.groups.empty()) o.groups.push_back(ObjGroup());
            ObjGroup &g = o.groups.back();
            for (size_t k = 0; k + 2 < vi.size(); ++k) {
                g.tv.insert(g.tv.end(), {vi[0], vi[k + 1], vi[k + 2]});
                g.tn.insert(g.tn.end(), {ni[0], ni[k + 1], ni[k + 2]});
            }
        }
    }
    for (const ObjGroup &g : o.groups) {
        for (int32_t x : g.tv)
            if (x >= (int32_t)o.verts.size()) return fail(TT_ERR_INVALID, path + ": vertex index out of range");
        if (!o.normals.empty())
            for (int32_t x : g.tn)
                if (x < 0 || x >= (int32_t)o.normals.size())
                    return fail(TT_ERR_INVALID, path + ": normal index out of range");
    }
    // Object::postProcess: centre the bounding box of the referenced vertices at the
    // origin and scale the largest coordinate magnitude to 1
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    bool any = false;
    for (const ObjGroup &g : o.groups)
        for (int32_t x : g.tv) {
            const V3 &p = o.verts[x];
            const float c[3] = {p.x, p.y, p.z};
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], c[k]);
                hi[k] = std::max(hi[k], c[k]);
            }
            any = true;
        }
    if (any) {
        const float ctr[3] = {(hi[0] + lo[0]) * 0.5f, (hi[1] + lo[1]) * 0.5f, (hi[2] + lo[2]) * 0.5f};
        float maxi = 0.0f;
        for (int k = 0; k < 3; ++k) {
            maxi = std::max(maxi, std::fabs(lo[k] - ctr[k]));
            maxi = std::max(maxi, std::fabs(hi[k] - ctr[k]));
        }
        const float sc = 1.0f / maxi;
        for (V3 &p : o.verts) {
            p.x = (p.x - ctr[0]) * sc;
            p.y = (p.y - ctr[1]) * sc;
            p.z = (p.z - ctr[2]) * sc;
        }
    }
    return TT_OK;
}

// ------------------------------------------------------------------ transforms
struct M4 {
    float m[4][4];
    static M4 ident() {
        M4 r{};
        for (int k = 0; k < 4; ++k) r.m[k][k] = 1.0f;
        return r;
    }
    M4 operator*(const M4 &b) const {
        M4 r{};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
                for (int k = 0; k < 4; ++k) s += m[i][k] * b.m[k][j];
                r.m[i][j] = s;
            }
        return r;
    }
    void apply(float x, float y, float z, float w, float out[3]) const {
        for (int i = 0; i < 3; ++i) out[i] = m[i][0] * x + m[i][1] * y + m[i][2] * z + m[i][3] * w;
    }
};

// Transform (Transform.cpp:13-34, PoseMath.h): M = S * T * (Ry * Rx * Rz), angles in
// degrees; applied as M (x, y, z, 1) to vertices and M (x, y, z, 0) to normals.
tt_status json_transform(const JVal &obj, M4 &out) {
    out = M4::ident();
    const JVal *t = obj.get("transform");
    if (!t) return TT_OK;
    float v[3];
    if (const JVal *s = t->get("scale")) {
        if (!jvec3(s, v)) return fail(TT_ERR_INVALID, "transform: bad scale");
        M4 S = M4::ident();
        S.m[0][0] = v[0], S.m[1][1] = v[1], S.m[2][2] = v[2];
        out = out * S;
    }
    if (const JVal *tr = t->get("translate")) {
        if (!jvec3(tr, v)) return fail(TT_ERR_INVALID, "transform: bad translate");
        M4 T = M4::ident();
        T.m[0][3] = v[0], T.m[1][3] = v[1], T.m[2][3] = v[2];
        out = out * T;
    }
    if (const JVal *o = t->get("orientation")) {
        if (!jvec3(o, v)) return fail(TT_ERR_INVALID, "transform: bad orientation");
        const float pi = (float)M_PI;
        const float rx = (v[0] / 180.0f) * pi, ry = (v[1] / 180.0f) * pi, rz = (v[2] / 180.0f) * pi;
        M4 X = M4::ident(), Y = M4::ident(), Z = M4::ident();
        X.m[1][1] = std::cos(rx), X.m[1][2] = -std::sin(rx), X.m[2][1] = std::sin(rx), X.m[2][2] = std::cos(rx);
        Y.m[0][0] = std::cos(ry), Y.m[0][2] = std::sin(ry), Y.m[2][0] = -std::sin(ry), Y.m[2][2] = std::cos(ry);
        Z.m[0][0] = std::cos(rz), Z.m[0][1] = -std::sin(rz), Z.m[1][0] = std::sin(rz), Z.m[1][1] = std::cos(rz);
        out = out * ((Y * X) * Z);
    }
    return TT_OK;
}

// Appends an OBJ object (transformed) to the scene mesh; material ids are offset
// by mat_base (the object's library follows the scene's materials). Objects
// without normals get one face normal per triangle.
void append_object(Mesh &scene, const Obj &o, const M4 &T, int32_t mat_base) {
    const int32_t vb = (int32_t)scene.verts.size(), nb = (int32_t)scene.normals.size();
    for (const V3 &p : o.verts) {
        float q[3];
        T.apply(p.x, p.y, p.z, 1.0f, q);
        scene.verts.push_back(V3{q[0], q[1], q[2]});
    }
    const bool has_n = !o.normals.empty();
    for (const V3 &n : o.normals) {
        float q[3];
        T.apply(n.x, n.y, n.z, 0.0f, q);
        const float l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        scene.normals.push_back(l > 0 ? V3{q[0] / l, q[1] / l, q[2] / l} : V3{0, 0, 1});
    }
    for (const ObjGroup &g : o.groups) {
        const int mi = g.material.empty() ? 0 : o.pool.index(g.material);
        for (size_t t = 0; t < g.tv.size() / 3; ++t) {
            for (int k = 0; k < 3; ++k) scene.tri_v.push_back(vb + g.tv[3 * t + k]);
            if (has_n) {
                for (int k = 0; k < 3; ++k) scene.tri_n.push_back(nb + g.tn[3 * t + k]);
            } else {
                const V3 &a = scene.verts[vb + g.tv[3 * t]], &b = scene.verts[vb + g.tv[3 * t + 1]],
                         &c = scene.verts[vb + g.tv[3 * t + 2]];
                const double ux = (double)b.x - a.x, uy = (double)b.y - a.y, uz = (double)b.z - a.z;
                const double vx = (double)c.x - a.x, vy = (double)c.y - a.y, vz = (double)c.z - a.z;
                double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
                const double l = std::sqrt(nx * nx + ny * ny + nz * nz);
                if (l > 0) nx /= l, ny /= l, nz /= l;
                else nx = 0, ny = 0, nz = 1;
                const int32_t ni = (int32_t)scene.normals.size();
                scene.normals.push_back(V3{(float)nx, (float)ny, (float)nz});
                for (int k = 0; k < 3; ++k) scene.tri_n.push_back(ni);
            }
            scene.tri_mat.push_back(mat_base + std::max(0, mi));
        }
    }
}

}  // namespace
}  // namespace tt

using namespace tt;

struct tt_config {
    tt_config_info info{};
    std::vector<tt_material> mats;
    std::vector<tt_geometry> geoms;
    Mesh mesh;
};

namespace {

struct Px {
    unsigned char c[4];
};

// RGBE pixel (r, g, b, e) -> floats (m / 256) * 2^(e - 128), HDRLoader.cpp:67-82
inline void rgbe_to_float(const unsigned char *p, float *out) {
    const float d = std::ldexp(1.0f, (int)p[3] - 128);
    for (int k = 0; k < 3; ++k) out[k] = ((float)p[k] / 256.0f) * d;
}

// Flat / old run-length scanline (HDRLoader.cpp:125-152): (1, 1, 1, n) repeats the
// previous pixel n << shift times, shift growing by 8 on consecutive repeats.
bool old_scanline(FILE *f, Px *px, int len, int start) {
    int shift = 0, x = start;
    while (x < len) {
        unsigned char p[4];
        for (int k = 0; k < 4; ++k) {
            int c = std::fgetc(f);
            if (c == EOF) return false;
            p[k] = (unsigned char)c;
        }
        if (p[0] == 1 && p[1] == 1 && p[2] == 1) {
            if (x == 0) return false;  // nothing to repeat
            for (long n = (long)p[3] << shift; n > 0 && x < len; --n, ++x) px[x] = px[x - 1];
            shift += 8;
        } else {
            std::memcpy(px[x++].c, p, 4);
            shift = 0;
        }
    }
    return true;
}

// Adaptive run-length scanline (HDRLoader.cpp:85-123): marker 2, 2, then the four
// components one after another, each as runs (count > 128: count - 128 copies of
// one byte) or literal spans.
bool read_scanline(FILE *f, Px *px, int len) {
    if (len < 8 || len > 0x7fff) return old_scanline(f, px, len, 0);
    int c0 = std::fgetc(f);
    if (c0 == EOF) return false;
    if (c0 != 2) {
        std::ungetc(c0, f);
        return old_scanline(f, px, len, 0);
    }
    int c1 = std::fgetc(f), c2 = std::fgetc(f), c3 = std::fgetc(f);
    if (c1 == EOF || c2 == EOF || c3 == EOF) return false;
    if (c1 != 2 || (c2 & 128)) {
        px[0].c[0] = 2, px[0].c[1] = (unsigned char)c1, px[0].c[2] = (unsigned char)c2, px[0].c[3] = (unsigned char)c3;
        return old_scanline(f, px, len, 1);
    }
    for (int comp = 0; comp < 4; ++comp) {
        for (int x = 0; x < len;) {
            int code = std::fgetc(f);
            if (code == EOF) return false;
            if (code > 128) {
                int n = code & 127;
                int v = std::fgetc(f);
                if (v == EOF || x + n > len) return false;
                while (n--) px[x++].c[comp] = (unsigned char)v;
            } else {
                int n = code;
                if (n == 0 || x + n > len) return false;
                while (n--) {
                    int v = std::fgetc(f);
                    if (v == EOF) return false;
                    px[x++].c[comp] = (unsigned char)v;
                }
            }
        }
    }
    return true;
}

void float_to_rgbe(const float *c, unsigned char *p) {
    const float v = std::max(c[0], std::max(c[1], c[2]));
    if (!(v > 1e-32f)) {
        p[0] = p[1] = p[2] = p[3] = 0;
        return;
    }
    int e;
    const float m = std::frexp(v, &e) * 256.0f / v;  // v = m' 2^e, m' in [0.5, 1)
    for (int k = 0; k < 3; ++k) p[k] = (unsigned char)std::min(255.0f, std::max(0.0f, c[k] * m));
    p[3] = (unsigned char)(e + 128);
}

}  // namespace

extern "C" {

tt_status tt_hdr_read(const char *path, int32_t *width, int32_t *height, float *rgb, int64_t capacity) {
    if (!path || !width || !height) return fail(TT_ERR_INVALID, "tt_hdr_read: null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(TT_ERR_IO, std::string("tt_hdr_read: cannot open ") + path);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(f, std::fclose);
    char magic[10];
    if (std::fread(magic, 1, 10, f) != 10 || std::memcmp(magic, "#?RADIANCE", 10) != 0)
        return fail(TT_ERR_IO, "tt_hdr_read: not a Radiance file");
    // header lines up to an empty line, then the resolution line "-Y h +X w"
    int c, prev = 0;
    while ((c = std::fgetc(f)) != EOF) {
        if (c == '\n' && prev == '\n') break;
        prev = c;
    }
    if (c == EOF) return fail(TT_ERR_IO, "tt_hdr_read: truncated header");
    char reso[256];
    if (!std::fgets(reso, sizeof(reso), f)) return fail(TT_ERR_IO, "tt_hdr_read: no resolution line");
    long h = 0, w = 0;
    if (std::sscanf(reso, "-Y %ld +X %ld", &h, &w) != 2 || w <= 0 || h <= 0 || w > (1 << 16) || h > (1 << 16))
        return fail(TT_ERR_IO, "tt_hdr_read: unsupported resolution line (only -Y h +X w)");
    *width = (int32_t)w;
    *height = (int32_t)h;
    if (!rgb) return TT_OK;
    if (capacity < (int64_t)w * h * 3) return fail(TT_ERR_INVALID, "tt_hdr_read: output buffer too small");
    std::vector<Px> px((size_t)w);
    for (long y = 0; y < h; ++y) {   // file order: top row first
        if (!read_scanline(f, px.data(), (int)w)) return fail(TT_ERR_IO, "tt_hdr_read: corrupt or short scanline");
        float *o = rgb + (size_t)y * w * 3;
        for (long x = 0; x < w; ++x) rgbe_to_float(px[x].c, o + 3 * x);
    }
    return TT_OK;
}

tt_status tt_hdr_write(const char *path, int32_t width, int32_t height, const float *rgb) {
    if (!path || !rgb || width <= 0 || height <= 0 || width > 0x7fff)
        return fail(TT_ERR_INVALID, "tt_hdr_write: bad arguments");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(TT_ERR_IO, std::string("tt_hdr_write: cannot open ") + path);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(f, std::fclose);
    std::fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", height, width);
    std::vector<unsigned char> comp((size_t)width);
    std::vector<Px> px((size_t)width);
    for (int32_t y = 0; y < height; ++y) {
        for (int32_t x = 0; x < width; ++x) float_to_rgbe(rgb + ((size_t)y * width + x) * 3, px[x].c);
        if (width < 8) {
            for (int32_t x = 0; x < width; ++x) std::fwrite(px[x].c, 1, 4, f);
            continue;
        }
        const unsigned char hdr[4] = {2, 2, (unsigned char)(width >> 8), (unsigned char)(width & 255)};
        std::fwrite(hdr, 1, 4, f);
        for (int k = 0; k < 4; ++k) {
            for (int32_t x = 0; x < width; ++x) comp[x] = px[x].c[k];
            // runs of >= 3 equal bytes as run codes, the rest as literal spans (<= 128)
            int32_t x = 0;
            while (x < width) {
                int32_t r = x + 1;
                while (r < width && r - x < 127 && comp[r] == comp[x]) ++r;
                if (r - x >= 3) {
                    std::fputc(128 + (r - x), f);
                    std::fputc(comp[x], f);
                    x = r;
                    continue;
                }
                int32_t l = x;
                while (l < width && l - x < 128) {
                    if (l + 2 < width && comp[l] == comp[l + 1] && comp[l] == comp[l + 2]) break;
                    ++l;
                }
                std::fputc(l - x, f);
                std::fwrite(&comp[x], 1, (size_t)(l - x), f);
                x = l;
            }
        }
    }
    return std::ferror(f) ? fail(TT_ERR_IO, "tt_hdr_write: write failed") : TT_OK;
}

tt_status tt_obj_load(const char *path, tt_mesh **mesh, tt_material *mats, int32_t mats_capacity, int32_t *num_mats) {
    if (!path || !mesh) return fail(TT_ERR_INVALID, "tt_obj_load: null argument");
    try {
        Obj o;
        tt_status st = load_obj(path, o);
        if (st != TT_OK) return st;
        auto *m = new tt_mesh();
        append_object(m->m, o, M4::ident(), 0);
        const int32_t nm = std::max<int32_t>(1, (int32_t)o.pool.mats.size());
        if (num_mats) *num_mats = nm;
        if (mats) {
            if (mats_capacity < nm) {
                delete m;
                return fail(TT_ERR_INVALID, "tt_obj_load: material buffer too small");
            }
            if (o.pool.mats.empty()) {
                tt_material d{};
                d.type = TT_MAT_DIFFUSE;
                d.albedo[0] = d.albedo[1] = d.albedo[2] = 0.8f;
                mats[0] = d;
            } else {
                std::copy(o.pool.mats.begin(), o.pool.mats.end(), mats);
            }
        }
        *mesh = m;
        return TT_OK;
    } catch (const std::exception &e) {
        return fail(TT_ERR_NOMEM, std::string("tt_obj_load: ") + e.what());
    }
}

tt_status tt_config_load(const char *path, const char *base_dir, tt_config **out) {
    if (!path || !out) return fail(TT_ERR_INVALID, "tt_config_load: null argument");
    try {
        const std::string base = base_dir && *base_dir ? std::string(base_dir) + "/" : std::string();
        JVal root;
        tt_status st = parse_json_file(path, root);
        if (st != TT_OK) return st;
        if (root.kind != JVal::OBJ) return fail(TT_ERR_INVALID, "tt_config_load: config must be a JSON object");
        auto cfg = std::make_unique<tt_config>();
        tt_config_info &in = cfg->info;
        auto need_str = [&](const char *k, char *dst) -> bool {
            const JVal *v = root.get(k);
            if (!v || v->kind != JVal::STR || v->str.size() >= 255) return false;
            std::strcpy(dst, v->str.c_str());
            return true;
        };
        const JVal *cam = root.get("camera");
        const JVal *spp = root.get("samples");
        const JVal *bvh_r = root.get("bvh-reload");
        const JVal *mat_r = root.get("material-reload");
        if (!cam || cam->kind != JVal::OBJ || !need_str("work-dir", in.work_dir) ||
            !need_str("hdr-file-name", in.hdr_file) || !need_str("material-file", in.material_file) || !spp ||
            spp->kind != JVal::NUM || !bvh_r || bvh_r->kind != JVal::BOOL || !mat_r || mat_r->kind != JVal::BOOL)
            return fail(TT_ERR_INVALID, "tt_config_load: missing or mistyped top-level field");
        in.samples = (int32_t)spp->num;
        in.bvh_reload = bvh_r->b;
        in.material_reload = mat_r->b;
        const JVal *w = cam->get("width"), *h = cam->get("height");
        if (!w || !h || w->kind != JVal::NUM || h->kind != JVal::NUM)
            return fail(TT_ERR_INVALID, "tt_config_load: camera width / height");
        in.width = (int32_t)w->num;
        in.height = (int32_t)h->num;
        // Camera (Camera.cpp:12-44): defaults position 0, yaw 0, pitch 0.3, focal 4,
        // aperture 0.04; radius fixed at 4; fovx 45 (main.cpp:38)
        in.position[0] = in.position[1] = in.position[2] = 0.0f;
        if (const JVal *p = cam->get("position"))
            if (!jvec3(p, in.position)) return fail(TT_ERR_INVALID, "tt_config_load: camera position");
        auto num_or = [&](const char *k, float d) {
            const JVal *v = cam->get(k);
            return v && v->kind == JVal::NUM ? (float)v->num : d;
        };
        in.yaw = num_or("yaw", 0.0f);
        in.pitch = num_or("pitch", 0.3f);
        in.focal = num_or("focal", 4.0f);
        in.aperture = num_or("aperture", 0.04f);
        in.radius = 4.0f;
        in.fovx = 45.0f;
        // materials: the master library first (key order), then each mesh object's
        // library in object order (Scene::add, Scene.cpp:7-37)
        Pool master;
        st = load_pool(base + in.material_file, master);
        if (st != TT_OK) return st;
        cfg->mats = master.mats;
        const JVal *objs = root.get("objects");
        if (objs && objs->kind != JVal::ARR) return fail(TT_ERR_INVALID, "tt_config_load: objects must be an array");
        if (objs)
            for (size_t k = 0; k < objs->arr.size(); ++k) {
                const JVal &o = objs->arr[k];
                const JVal *type = o.get("type");
                if (!type || type->kind != JVal::STR) return fail(TT_ERR_INVALID, "tt_config_load: object without type");
                M4 T;
                st = json_transform(o, T);
                if (st != TT_OK) return st;
                if (type->str == "Mesh") {
                    const JVal *file = o.get("file");
                    if (!file || file->kind != JVal::STR) return fail(TT_ERR_INVALID, "tt_config_load: Mesh without file");
                    Obj obj;
                    st = load_obj(base + file->str, obj);
                    if (st != TT_OK) return st;
                    append_object(cfg->mesh, obj, T, (int32_t)cfg->mats.size());
                    if (obj.pool.mats.empty()) {
                        tt_material d{};
                        d.type = TT_MAT_DIFFUSE;
                        d.albedo[0] = d.albedo[1] = d.albedo[2] = 0.8f;
                        cfg->mats.push_back(d);
                    } else {
                        cfg->mats.insert(cfg->mats.end(), obj.pool.mats.begin(), obj.pool.mats.end());
                    }
                } else if (type->str == "Plane" || type->str == "Sphere") {
                    const JVal *mn = o.get("material");
                    if (!mn || mn->kind != JVal::STR) return fail(TT_ERR_INVALID, "tt_config_load: geometry without material");
                    const int mi = master.index(mn->str);
                    if (mi < 0) return fail(TT_ERR_INVALID, "tt_config_load: material " + mn->str + " not found");
                    tt_geometry g{};
                    g.material = mi;
                    if (type->str == "Plane") {
                        g.type = TT_GEOM_PLANE;
                        if (!jvec3(o.get("p0"), g.a) || !jvec3(o.get("p1"), g.b) || !jvec3(o.get("p2"), g.c))
                            return fail(TT_ERR_INVALID, "tt_config_load: Plane needs p0, p1, p2");
                    } else {
                        g.type = TT_GEOM_SPHERE;
                        const JVal *r = o.get("radius");
                        if (!jvec3(o.get("position"), g.a) || !r || r->kind != JVal::NUM)
                            return fail(TT_ERR_INVALID, "tt_config_load: Sphere needs position, radius");
                        g.b[0] = (float)r->num;
                    }
                    cfg->geoms.push_back(g);
                } else {
                    return fail(TT_ERR_INVALID, "tt_config_load: unknown object type " + type->str);
                }
            }
        in.num_materials = (int32_t)cfg->mats.size();
        in.num_geometries = (int32_t)cfg->geoms.size();
        in.num_tris = cfg->mesh.ntris();
        *out = cfg.release();
        return TT_OK;
    } catch (const std::exception &e) {
        return fail(TT_ERR_NOMEM, std::string("tt_config_load: ") + e.what());
    }
}

void tt_config_destroy(tt_config *c) { delete c; }

tt_status tt_config_get(const tt_config *c, tt_config_info *info, tt_material *mats, tt_geometry *geoms, tt_mesh **mesh) {
    if (!c) return fail(TT_ERR_INVALID, "tt_config_get: null config");
    if (info) *info = c->info;
    if (mats) std::copy(c->mats.begin(), c->mats.end(), mats);
    if (geoms) std::copy(c->geoms.begin(), c->geoms.end(), geoms);
    if (mesh) {
        if (c->mesh.ntris() == 0) {
            *mesh = nullptr;
        } else {
            auto *m = new (std::nothrow) tt_mesh();
            if (!m) return fail(TT_ERR_NOMEM, "tt_config_get: host allocation failed");
            m->m = c->mesh;
            *mesh = m;
        }
    }
    return TT_OK;
}

}  // extern "C"
