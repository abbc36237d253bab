// Scene files: the TOML scene format of the reference's example runner, the regions it
// describes and the staggered particle lattice generated in them -- everything
// FluidSim::new does on the host before the first dispatch.  Host-only code of libarp.so
// (no HIP calls except through arp_scene_apply, which calls the public entry points).
//
//   examples/test_mods/config.rs:298-474   Config::parse (keys, defaults, material order)
//   examples/test_mods/config.rs:110-257   regions of a table (CSG, transforms, models)
//   examples/test_mods/mesh.rs:1-140       Mesh / Surface regions of an STL model
//   src/soft_body/material_region/region.rs:11-272   AABB, AABE, Transformed, CSG
//   src/soft_body/material_region/mod.rs:56-155      MaterialRegion::gen_particles
//   src/soft_body/mod.rs:58-111            FluidSim::new: materials, boundary split, matrix
//   src/soft_body/particle_state/particles.rs:488-507  init_solid_particles
//
// Arithmetic follows the reference's f32 (and, where it uses them, f64) operations in the
// same order, built without floating-point contraction, so the particle positions, masses
// and interaction radii are the reference's bit for bit.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/arp.h"
#include "toml.h"

using arp_toml::Kind;
using arp_toml::Value;

namespace {

constexpr double kPi = 3.14159265358979323846;

struct SceneError {
    std::string msg;
};

[[noreturn]] void bad(const char* fmt, ...) {
    char b[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(b, sizeof b, fmt, ap);
    va_end(ap);
    throw SceneError{b};
}

// ---- regions --------------------------------------------------------------------------
struct Box {
    float min[4] = {0, 0, 0, 0};
    float dim[4] = {0, 0, 0, 0};
    void center(float c[4]) const {
        for (int i = 0; i < 4; i++) c[i] = min[i] + 0.5f * dim[i];
    }
    void max(float m[4]) const {
        for (int i = 0; i < 4; i++) m[i] = min[i] + dim[i];
    }
    static Box from_min_max(const float mn[4], const float mx[4]) {
        Box b;
        for (int i = 0; i < 4; i++) {
            b.min[i] = mn[i];
            b.dim[i] = mx[i] - mn[i];
        }
        return b;
    }
};

struct Region {
    virtual ~Region() = default;
    virtual Box bound() const = 0;
    virtual bool contains(const float p[4]) const = 0;
};
using RegionP = std::shared_ptr<const Region>;

// AABB (region.rs:62-72): p[i] >= min[i] and p[i] - min[i] <= dim[i] on all four axes
struct AabbRegion : Region {
    Box b;
    explicit AabbRegion(const Box& x) : b(x) {}
    Box bound() const override { return b; }
    bool contains(const float p[4]) const override {
        for (int i = 0; i < 4; i++)
            if (p[i] < b.min[i] || p[i] - b.min[i] > b.dim[i]) return false;
        return true;
    }
};

// AABE (region.rs:74-107): the ellipsoid test multiplied through by the radii, in f64
struct AabeRegion : Region {
    float c[4], r[4];
    AabeRegion(const float cc[4], const float rr[4]) {
        for (int i = 0; i < 4; i++) {
            c[i] = cc[i];
            r[i] = rr[i];
        }
    }
    Box bound() const override {
        Box b;
        for (int i = 0; i < 4; i++) {
            b.min[i] = c[i] - r[i];
            b.dim[i] = 2.0f * r[i];
        }
        return b;
    }
    bool contains(const float p[4]) const override {
        double lhs = 0.0;
        for (int i = 0; i < 4; i++) {
            const double d = (double)(p[i] - c[i]);
            const double ri = (double)r[i];
            if (ri > 0.0) {
                double coeff = 1.0;
                for (int j = 0; j < 4; j++)
                    if (r[j] > 0.0f && j != i) coeff = coeff * (double)r[j];
                lhs = lhs + d * d * coeff * coeff;
            }
        }
        double rhs = 1.0;
        for (int i = 0; i < 4; i++)
            if (r[i] > 0.0f) rhs = rhs * (double)r[i];
        return lhs <= rhs * rhs;
    }
};

struct DifferenceRegion : Region {
    RegionP a, b;
    DifferenceRegion(RegionP x, RegionP y) : a(std::move(x)), b(std::move(y)) {}
    Box bound() const override { return a->bound(); }
    bool contains(const float p[4]) const override { return a->contains(p) && !b->contains(p); }
};

struct UnionRegion : Region {
    RegionP a, b;
    UnionRegion(RegionP x, RegionP y) : a(std::move(x)), b(std::move(y)) {}
    Box bound() const override {
        const Box l = a->bound(), r = b->bound();
        float m1[4], m2[4], mn[4], mx[4];
        l.max(m1);
        r.max(m2);
        for (int i = 0; i < 4; i++) {
            mn[i] = std::fmin(l.min[i], r.min[i]);
            mx[i] = std::fmax(m1[i], m2[i]);
        }
        return Box::from_min_max(mn, mx);
    }
    bool contains(const float p[4]) const override { return a->contains(p) || b->contains(p); }
};

struct IntersectionRegion : Region {
    RegionP a, b;
    IntersectionRegion(RegionP x, RegionP y) : a(std::move(x)), b(std::move(y)) {}
    Box bound() const override {
        const Box l = a->bound(), r = b->bound();
        float m1[4], m2[4];
        l.max(m1);
        r.max(m2);
        Box o;
        for (int i = 0; i < 4; i++) {
            o.min[i] = std::fmax(l.min[i], r.min[i]);
            o.dim[i] = std::fmin(m1[i], m2[i]) - o.min[i];
        }
        return o;
    }
    bool contains(const float p[4]) const override { return a->contains(p) && b->contains(p); }
};

// AABB::border / AABE::border (region.rs:33-58): the shell of depth b
RegionP aabb_border(const Box& x, float b) {
    float c[4];
    x.center(c);
    Box in;
    for (int i = 0; i < 4; i++) {
        in.min[i] = std::fmin(x.min[i] + b, c[i]);
        in.dim[i] = std::fmax(x.dim[i] - 2.0f * b, 0.0f);
    }
    return std::make_shared<DifferenceRegion>(std::make_shared<AabbRegion>(x), std::make_shared<AabbRegion>(in));
}

RegionP aabe_border(const float c[4], const float r[4], float b) {
    float in[4];
    for (int i = 0; i < 4; i++) in[i] = std::fmax(r[i] - b, 0.0f);
    return std::make_shared<DifferenceRegion>(std::make_shared<AabeRegion>(c, r), std::make_shared<AabeRegion>(c, in));
}

// Transformed(region, M, t) (region.rs:126-225): y = M x + t with M[i][j] = m[j][i] (m holds
// the matrix by columns, as a GLSL mat4); contains() solves M x = p - t by the reference's
// Gauss-Jordan elimination, operation for operation.
struct TransformedRegion : Region {
    RegionP base;
    float m[4][4];  // m[col][row]
    float t[4];
    TransformedRegion(RegionP r, const float mm[4][4], const float tt[4]) : base(std::move(r)) {
        std::memcpy(m, mm, sizeof m);
        std::memcpy(t, tt, sizeof t);
    }
    Box bound() const override {
        const Box b = base->bound();
        float mn[4], mx[4];
        for (int i = 0; i < 4; i++) {
            mn[i] = t[i];
            mx[i] = t[i];
        }
        for (int i = 0; i < 4; i++) {
            float a_min = 0.0f;
            for (int j = 0; j < 4; j++) a_min += m[j][i] * b.min[j];
            mn[i] += a_min;
            mx[i] += a_min;
            for (int j = 0; j < 4; j++) mn[i] += std::fmin(m[j][i] * b.dim[j], 0.0f);
            for (int j = 0; j < 4; j++) mx[i] += std::fmax(m[j][i] * b.dim[j], 0.0f);
        }
        return Box::from_min_max(mn, mx);
    }
    static void row_mul(float a[4][4], float v[4], int row, int col, float f) {
        for (int k = col; k < 4; k++) a[k][row] *= f;
        v[row] *= f;
    }
    static void row_sum(float a[4][4], float v[4], int row, int col, float f, int dest) {
        for (int k = col; k < 4; k++) a[k][dest] += f * a[k][row];
        v[dest] += f * v[row];
    }
    static void swap_rows(float a[4][4], float v[4], int r1, int r2, int col) {
        for (int k = col; k < 4; k++) std::swap(a[k][r1], a[k][r2]);
        std::swap(v[r1], v[r2]);
    }
    bool contains(const float p[4]) const override {
        float v[4], a[4][4];
        for (int i = 0; i < 4; i++) v[i] = p[i] - t[i];
        std::memcpy(a, m, sizeof a);
        int i = 0, j = 0;
        while (i < 4 && j < 4) {
            if (a[j][i] != 0.0f) {
                const float pivot = a[j][i];
                if (pivot != 1.0f) row_mul(a, v, i, j, 1.0f / pivot);
                for (int k = 0; k < 4; k++)
                    if (k != i && a[j][k] != 0.0f) {
                        const float pivot2 = -a[j][k];
                        row_sum(a, v, i, j, pivot2, k);
                    }
                i++;
                j++;
            } else {
                bool cont = false;
                for (int k = i + 1; k < 4; k++)
                    if (a[j][k] != 0.0f) {
                        swap_rows(a, v, i, k, j);
                        cont = true;
                    }
                if (!cont) j++;
            }
        }
        return base->contains(v);
    }
};

// ---- STL models (mesh.rs) ---------------------------------------------------------------
struct TriMesh {
    std::vector<float> verts;     // 3 per vertex, first-appearance order, duplicates merged
    std::vector<uint32_t> faces;  // 3 vertex indices per face
    std::vector<float> normals;   // 3 per face, as stored in the file
};

inline float signum(float x) { return std::isnan(x) ? x : std::copysign(1.0f, x); }

Box point_bound(const TriMesh& m) {
    if (m.verts.empty()) bad("STL model has no triangles");
    float mn[4] = {m.verts[0], m.verts[1], m.verts[2], 0.0f};
    float mx[4] = {mn[0], mn[1], mn[2], 0.0f};
    for (size_t v = 0; v < m.verts.size() / 3; v++)
        for (int i = 0; i < 3; i++) {
            mn[i] = std::fmin(m.verts[3 * v + i], mn[i]);
            mx[i] = std::fmax(m.verts[3 * v + i], mx[i]);
        }
    return Box::from_min_max(mn, mx);
}

// Mesh (mesh.rs:24-73): parity of the faces at or above the point whose xy projection
// holds it
struct MeshRegion : Region {
    std::shared_ptr<const TriMesh> mesh;
    explicit MeshRegion(std::shared_ptr<const TriMesh> m) : mesh(std::move(m)) {}
    Box bound() const override { return point_bound(*mesh); }
    bool contains(const float p[4]) const override {
        if (p[3] != 0.0f) return false;
        const float* V = mesh->verts.data();
        bool r = false;
        const size_t nf = mesh->faces.size() / 3;
        for (size_t f = 0; f < nf; f++) {
            const float* v0 = V + 3 * (size_t)mesh->faces[3 * f];
            const float* v1 = V + 3 * (size_t)mesh->faces[3 * f + 1];
            const float* v2 = V + 3 * (size_t)mesh->faces[3 * f + 2];
            if (!(v0[2] >= p[2] || v1[2] >= p[2] || v2[2] >= p[2])) continue;
            const float s1x = v1[0] - v0[0], s1y = v1[1] - v0[1];
            const float s2x = v2[0] - v0[0], s2y = v2[1] - v0[1];
            const float dx = p[0] - v0[0], dy = p[1] - v0[1];
            float a = s1x * s2y - s1y * s2x;
            const float sg = signum(a);
            const float s = sg * (dx * s2y - dy * s2x);
            const float t = sg * (s1x * dy - s1y * dx);
            a = std::fabs(a);
            const bool hit = s >= 0.0f && s <= a && t >= 0.0f && t <= a && s + t <= a;
            r = r != hit;
        }
        return r;
    }
};

inline float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3(const float* a, const float* b, float* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
inline void sub3(const float* a, const float* b, float* o) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}

// Surface (mesh.rs:76-138): within `border` of a vertex, or of a face inside its prism
struct SurfaceRegion : Region {
    std::shared_ptr<const TriMesh> mesh;
    float border;
    SurfaceRegion(std::shared_ptr<const TriMesh> m, float b) : mesh(std::move(m)), border(b) {}
    Box bound() const override {
        Box b = point_bound(*mesh);
        for (int i = 0; i < 3; i++) {
            b.min[i] -= border;
            b.dim[i] += border * 2.0f;
        }
        return b;
    }
    bool contains(const float p[4]) const override {
        const float* V = mesh->verts.data();
        const float bb = border * border;
        for (size_t v = 0; v < mesh->verts.size() / 3; v++) {
            float d[3];
            sub3(p, V + 3 * v, d);
            if (dot3(d, d) <= bb) return true;
        }
        const size_t nf = mesh->faces.size() / 3;
        for (size_t f = 0; f < nf; f++) {
            const float* nrm = mesh->normals.data() + 3 * f;
            const float* a = V + 3 * (size_t)mesh->faces[3 * f];
            const float* b = V + 3 * (size_t)mesh->faces[3 * f + 1];
            const float* c = V + 3 * (size_t)mesh->faces[3 * f + 2];
            float d[3];
            sub3(p, a, d);
            if (std::fabs(dot3(d, nrm)) <= std::sqrt(dot3(nrm, nrm)) * border) {
                float s1[3], s2[3], s3[3], s4[3], d2[3], n1[3], n2[3], n3[3];
                sub3(b, a, s1);
                sub3(c, a, s2);
                sub3(c, b, s3);
                sub3(a, b, s4);
                sub3(p, b, d2);
                cross3(s1, nrm, n1);
                cross3(s2, nrm, n2);
                cross3(s3, nrm, n3);
                if (signum(dot3(d, n1)) == signum(dot3(s2, n1)) && signum(dot3(d, n2)) == signum(dot3(s1, n2)) &&
                    signum(dot3(d2, n3)) == signum(dot3(s4, n3)))
                    return true;
            }
        }
        return false;
    }
};

// STL reader (the stl_io 0.4 crate's read_stl): ASCII when the file opens with "solid" and
// its next line with "facet" or "endsolid", binary otherwise; vertices merged by bit pattern
// in order of first appearance.
std::shared_ptr<const TriMesh> read_stl(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) bad("cannot open STL model '%s'", path.c_str());
    std::string data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    auto m = std::make_shared<TriMesh>();
    std::unordered_map<uint64_t, std::vector<std::pair<std::array<uint32_t, 3>, uint32_t>>> seen;
    auto vertex = [&](const float* x) -> uint32_t {
        std::array<uint32_t, 3> bits;
        std::memcpy(bits.data(), x, 12);
        const uint64_t h = ((uint64_t)bits[0] * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)bits[1] << 21) ^
                           ((uint64_t)bits[2] * 0xC2B2AE3D27D4EB4Full);
        auto& bucket = seen[h];
        for (auto& e : bucket)
            if (e.first == bits) return e.second;
        const uint32_t idx = (uint32_t)(m->verts.size() / 3);
        m->verts.insert(m->verts.end(), x, x + 3);
        bucket.push_back({bits, idx});
        return idx;
    };
    auto add_face = [&](const float* nrm, const float* v) {
        m->normals.insert(m->normals.end(), nrm, nrm + 3);
        for (int k = 0; k < 3; k++) m->faces.push_back(vertex(v + 3 * k));
    };
    bool ascii = false;
    if (data.compare(0, 5, "solid") == 0) {
        size_t e = data.find('\n');
        if (e != std::string::npos) {
            size_t q = e + 1;
            while (q < data.size() && std::isspace((unsigned char)data[q])) q++;
            ascii = data.compare(q, 5, "facet") == 0 || data.compare(q, 8, "endsolid") == 0;
        }
    }
    if (ascii) {
        const char* s = data.c_str();
        const char* end = s + data.size();
        auto word = [&](std::string& w) -> bool {
            while (s < end && std::isspace((unsigned char)*s)) s++;
            const char* b = s;
            while (s < end && !std::isspace((unsigned char)*s)) s++;
            w.assign(b, s - b);
            return !w.empty();
        };
        auto num = [&]() -> float {
            std::string w;
            if (!word(w)) bad("STL model '%s': truncated", path.c_str());
            char* e = nullptr;
            const float x = std::strtof(w.c_str(), &e);
            if (*e != '\0') bad("STL model '%s': bad number '%s'", path.c_str(), w.c_str());
            return x;
        };
        auto expect = [&](const char* lit) {
            std::string w;
            if (!word(w) || w != lit) bad("STL model '%s': expected '%s'", path.c_str(), lit);
        };
        std::string w;
        word(w);  // "solid"
        // the rest of the solid line is its name
        while (s < end && *s != '\n') s++;
        while (word(w)) {
            if (w == "endsolid") break;
            if (w != "facet") bad("STL model '%s': expected 'facet'", path.c_str());
            expect("normal");
            float nrm[3], v[9];
            for (int k = 0; k < 3; k++) nrm[k] = num();
            expect("outer");
            expect("loop");
            for (int q = 0; q < 3; q++) {
                expect("vertex");
                for (int k = 0; k < 3; k++) v[3 * q + k] = num();
            }
            expect("endloop");
            expect("endfacet");
            add_face(nrm, v);
        }
    } else {
        if (data.size() < 84) bad("STL model '%s': truncated header", path.c_str());
        uint32_t nt;
        std::memcpy(&nt, data.data() + 80, 4);
        if (data.size() < 84 + 50 * (size_t)nt) bad("STL model '%s': truncated (%u triangles)", path.c_str(), nt);
        for (uint32_t t = 0; t < nt; t++) {
            float rec[12];
            std::memcpy(rec, data.data() + 84 + 50 * (size_t)t, 48);
            add_face(rec, rec + 3);
        }
    }
    if (m->faces.empty()) bad("STL model '%s' has no triangles", path.c_str());
    return m;
}

// ---- typed access to the TOML tree (config.rs:59-117) ----------------------------------
struct Ctx {
    std::string where;
};

double as_float_or(const Value& t, const char* k, double def, const std::string& where) {
    const Value* v = t.get(k);
    if (!v) return def;
    if (!v->is_float()) bad("%s: '%s' must be a float, not %s", where.c_str(), k, arp_toml::kind_name(v->kind));
    return v->f;
}

int64_t as_int_or(const Value& t, const char* k, int64_t def, const std::string& where) {
    const Value* v = t.get(k);
    if (!v) return def;
    if (!v->is_int()) bad("%s: '%s' must be an integer, not %s", where.c_str(), k, arp_toml::kind_name(v->kind));
    return v->i;
}

bool as_bool_or(const Value& t, const char* k, bool def, const std::string& where) {
    const Value* v = t.get(k);
    if (!v) return def;
    if (!v->is_bool()) bad("%s: '%s' must be a boolean", where.c_str(), k);
    return v->b;
}

std::string as_str_or(const Value& t, const char* k, const std::string& def, const std::string& where) {
    const Value* v = t.get(k);
    if (!v) return def;
    if (!v->is_str()) bad("%s: '%s' must be a string", where.c_str(), k);
    return v->str;
}

// to_vec4: the first (up to) four entries as f32, the rest 0
void to_vec4(const Value& a, float out[4], const char* k, const std::string& where) {
    for (int i = 0; i < 4; i++) out[i] = 0.0f;
    for (size_t i = 0; i < 4 && i < a.arr.size(); i++) {
        if (!a.arr[i]->is_float()) bad("%s: entries of '%s' must be floats", where.c_str(), k);
        out[i] = (float)a.arr[i]->f;
    }
}

void as_vec4_or(const Value& t, const char* k, const float def[4], float out[4], const std::string& where) {
    const Value* v = t.get(k);
    if (!v) {
        for (int i = 0; i < 4; i++) out[i] = def[i];
        return;
    }
    if (!v->is_array()) bad("%s: '%s' must be an array", where.c_str(), k);
    to_vec4(*v, out, k, where);
}

// ---- regions of a table (config.rs:110-257) ----------------------------------------------
struct RegionParser {
    std::string base_dir;
    std::unordered_map<std::string, std::shared_ptr<const TriMesh>> models;

    std::shared_ptr<const TriMesh> model(const std::string& path) {
        std::string full = path;
        if (!base_dir.empty() && !path.empty() && path[0] != '/') full = base_dir + "/" + path;
        auto it = models.find(full);
        if (it != models.end()) return it->second;
        auto m = read_stl(full);
        models[full] = m;
        return m;
    }

    RegionP fold(const Value& arr, int op, const std::string& where) {
        if (arr.arr.empty()) bad("%s: empty region list", where.c_str());
        RegionP acc;
        for (size_t k = 0; k < arr.arr.size(); k++) {
            if (!arr.arr[k]->is_table()) bad("%s: region list entries must be tables", where.c_str());
            RegionP r = region(*arr.arr[k], where);
            if (!acc) {
                acc = r;
            } else if (op == 0) {
                acc = std::make_shared<UnionRegion>(acc, r);
            } else if (op == 1) {
                acc = std::make_shared<DifferenceRegion>(acc, r);
            } else {
                acc = std::make_shared<IntersectionRegion>(acc, r);
            }
        }
        return acc;
    }

    // parse_region_from_mat: union / difference / intersection arrays, or a region table
    RegionP from_mat(const Value& t, const std::string& where) {
        const Value* v;
        if ((v = t.get("union")) && v->is_array()) return fold(*v, 0, where);
        if ((v = t.get("difference")) && v->is_array()) return fold(*v, 1, where);
        if ((v = t.get("intersection")) && v->is_array()) return fold(*v, 2, where);
        if ((v = t.get("region")) && v->is_table()) return region(*v, where);
        bad("%s: no table named region and no union, difference or intersection array", where.c_str());
    }

    // parse_region: a base region, wrapped in a transform when the table has one
    RegionP region(const Value& t, const std::string& where) {
        RegionP base = base_region(t, where);
        const Value* tr = t.get("translation");
        const Value* sc = t.get("scale");
        const Value* ro = t.get("rotation");
        if (!tr && !sc && !ro) return base;
        const float z4[4] = {0, 0, 0, 0};
        float trans[4];
        as_vec4_or(t, "translation", z4, trans, where);
        float rot[4];
        if (ro && ro->is_float()) {
            rot[0] = 0.0f;
            rot[1] = 0.0f;
            rot[2] = 1.0f;
            rot[3] = (float)ro->f;
        } else {
            const float d[4] = {0, 0, 1, 0};
            as_vec4_or(t, "rotation", d, rot, where);
        }
        float scale[4] = {1, 1, 1, 1};
        if (sc && sc->is_float()) {
            for (int i = 0; i < 4; i++) scale[i] = (float)sc->f;
        } else if (sc && sc->is_array()) {
            for (size_t i = 0; i < sc->arr.size() && i < 4; i++) {
                if (!sc->arr[i]->is_float()) bad("%s: entries of 'scale' must be floats", where.c_str());
                scale[i] = (float)sc->arr[i]->f;
            }
        }
        // f32::to_radians multiplies by PI/180 formed in f32
        const float rad = rot[3] * ((float)kPi / 180.0f);
        const float c = std::cos(rad);
        const float s = std::sin(rad);
        const float u_l = std::sqrt(rot[0] * rot[0] + rot[1] * rot[1] + rot[2] * rot[2]);
        const float u[3] = {rot[0] / u_l, rot[1] / u_l, rot[2] / u_l};
        // the reference's rows as written (row 1's first entry included), row i scaled
        // by scale[i]; each row is a column of the GLSL mat4
        float m[4][4] = {
            {c + u[0] * u[0] * (1.0f - c), u[0] * u[1] * (1.0f - c) + u[2] * s, u[0] * u[2] * (1.0f - c) - u[1] * s,
             0.0f},
            {u[1] * u[2] * (1.0f - c) - u[1] * s, c + u[1] * u[1] * (1.0f - c), u[1] * u[2] * (1.0f - c) + u[0] * s,
             0.0f},
            {u[0] * u[2] * (1.0f - c) + u[1] * s, u[1] * u[2] * (1.0f - c) - u[0] * s, c + u[2] * u[2] * (1.0f - c),
             0.0f},
            {0.0f, 0.0f, 0.0f, 1.0f}};
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) m[i][j] *= scale[i];
        return std::make_shared<TransformedRegion>(base, m, trans);
    }

    // parse_base: a model, a box (dim), an ellipsoid (radii / diameters), or a nested CSG
    RegionP base_region(const Value& t, const std::string& where) {
        const Value* mdl = t.get("model");
        if (mdl && mdl->is_str()) {
            const Value* b = t.get("border");
            if (b && b->is_float()) return std::make_shared<SurfaceRegion>(model(mdl->str), (float)b->f);
            return std::make_shared<MeshRegion>(model(mdl->str));
        }
        bool boxed;
        float dim[4], mn[4] = {0, 0, 0, 0};
        const Value* bd = t.get("border");
        if (bd && !bd->is_float()) bad("%s: 'border' must be a float", where.c_str());
        const Value* v;
        if ((v = t.get("dim")) && v->is_array()) {
            boxed = true;
            to_vec4(*v, dim, "dim", where);
        } else if ((v = t.get("radii")) && v->is_array()) {
            boxed = false;
            to_vec4(*v, dim, "radii", where);
            for (int i = 0; i < 4; i++) dim[i] *= 2.0f;
        } else if ((v = t.get("diameters")) && v->is_array()) {
            boxed = false;
            to_vec4(*v, dim, "diameters", where);
        } else {
            return from_mat(t, where);
        }
        if ((v = t.get("min")) && v->is_array()) {
            to_vec4(*v, mn, "min", where);
        } else if ((v = t.get("center")) && v->is_array()) {
            to_vec4(*v, mn, "center", where);
            for (int i = 0; i < 4; i++) mn[i] -= 0.5f * dim[i];
        }
        if (boxed) {
            Box b;
            std::memcpy(b.min, mn, sizeof mn);
            std::memcpy(b.dim, dim, sizeof dim);
            if (bd) return aabb_border(b, (float)bd->f);
            return std::make_shared<AabbRegion>(b);
        }
        for (int i = 0; i < 4; i++) {
            dim[i] *= 0.5f;
            mn[i] += dim[i];
        }
        if (bd) return aabe_border(mn, dim, (float)bd->f);
        return std::make_shared<AabeRegion>(mn, dim);
    }
};

// ---- objects ----------------------------------------------------------------------------
struct Object {
    std::string name;
    float colors[3][4];
    RegionP region;
    float packing = 0.5f;
    ArpMaterial mat{};
    bool moving = false;  // velocity field of an [objects.*] table
    float vel[4] = {0, 0, 0, 0};
    float ang[6] = {0, 0, 0, 0, 0, 0};
    float center[4] = {0, 0, 0, 0};
    bool relative = false;  // on_click = "Relative"
};

}  // namespace

struct arp_scene_file {
    std::string title = "Fluid Test", integrator = "verlet";
    bool lighting = false;
    uint32_t subticks = 1;
    float h = 1.0f / 64.0f, dt = 0.01f, alpha = 50.0f, g = 1.0f, render_h = 0.0f;
    float min[4] = {-1, -1, 0, 0}, dim[4] = {2, 2, 0, 0};
    float light_pos[3], ambient, diffuse, specular, light_color[3];
    std::vector<double> view_pos;
    double view_scale = 1.0, view_angle = 0.0;
    std::vector<Object> objects, on_click;
    std::vector<ArpMatInteraction> inter_cfg;  // nobj x nobj configured pairs
    std::vector<uint8_t> inter_present;
    // generated
    bool generated = false;
    std::vector<ArpParticle> particles, boundary;
    std::vector<ArpSolidParticle> solids;
    std::vector<ArpMaterial> mats;
    std::vector<ArpMatInteraction> inter;
    std::vector<uint32_t> first, count, tested;
    std::string err;
};

namespace {

void get_colors(const Value& t, float c[3][4], const std::string& where) {
    const float d[4] = {0, 0, 0, 1};
    as_vec4_or(t, "color", d, c[0], where);
    as_vec4_or(t, "color_low_density", c[0], c[1], where);
    as_vec4_or(t, "color_high_density", c[0], c[2], where);
}

// get_packing: packing, else spacing / h (f64), else 0.5
float get_packing(const Value& t, float h, const std::string& where) {
    const double sp = as_float_or(t, "spacing", 0.5 * (double)h, where);
    return (float)as_float_or(t, "packing", sp / (double)h, where);
}

// parse_interation (config.rs:259-282)
ArpMatInteraction parse_interaction(const Value& t, double h, const std::string& where) {
    ArpMatInteraction m{};
    m.strength = (float)as_float_or(t, "contact_strength", 10000.0, where);
    m.radius = (float)as_float_or(t, "contact_radius", as_float_or(t, "contact_factor", 2.0, where) * h, where);
    m.friction = (float)as_float_or(t, "friction", 0.0, where);
    m.dampening = (float)as_float_or(t, "dampening", 0.0, where);
    const Value* pv = t.get("potential");
    uint32_t pot = 2;  // Linear, also for any unknown name
    if (pv) {
        if (!pv->is_str()) bad("%s: 'potential' must be a string", where.c_str());
        const std::string& s = pv->str;
        if (s == "Zero" || s == "zero") pot = 0;
        else if (s == "Constant" || s == "constant") pot = 1;
        else if (s == "Linear" || s == "linear") pot = 2;
        else if (s == "Tait" || s == "tait") pot = 3;
        else if (s == "Hertz" || s == "hertz") pot = 5;
        else if (s == "LennardJones" || s == "Lennard-Jones" || s == "Lennard Jones" || s == "lennard jones" ||
                 s == "lennard-jones")
            pot = 4;
    }
    m.potential = pot;
    return m;
}

using InterMap = std::vector<std::pair<std::pair<std::string, std::string>, ArpMatInteraction>>;

void put_inter(InterMap& map, const std::string& a, const std::string& b, const ArpMatInteraction& m) {
    for (auto& e : map)
        if (e.first.first == a && e.first.second == b) {
            e.second = m;
            return;
        }
    map.push_back({{a, b}, m});
}

const ArpMatInteraction* find_inter(const InterMap& map, const std::string& a, const std::string& b) {
    for (auto& e : map)
        if (e.first.first == a && e.first.second == b) return &e.second;
    return nullptr;
}

// get_interactions: every table-valued key other than region / union / intersection /
// difference names the material it configures
void get_interactions(const std::string& name, const Value& obj, double h, InterMap& map, const std::string& where) {
    for (auto& kv : obj.tab) {
        if (!kv.second->is_table()) continue;
        const std::string& k = kv.first;
        if (k == "region" || k == "union" || k == "intersection" || k == "difference") continue;
        put_inter(map, name, k, parse_interaction(*kv.second, h, where + "." + k));
    }
}

void parse_config(arp_scene_file& S, const Value& root, RegionParser& rp) {
    const std::string W = "scene";
    S.title = as_str_or(root, "name", "Fluid Test", W);
    S.lighting = as_bool_or(root, "lighting", root.get("light_pos") != nullptr, W);
    const int64_t st = as_int_or(root, "subticks", 1, W);
    if (st < 0 || st > 0xFFFFFFFFll) bad("scene: 'subticks' out of range");
    S.subticks = (uint32_t)st;
    S.h = (float)as_float_or(root, "kernel_radius", 1.0 / 64.0, W);
    S.dt = (float)as_float_or(root, "time_step", 0.01, W);
    S.alpha = (float)as_float_or(root, "artificial_viscocity", 50.0, W);
    S.g = (float)as_float_or(root, "gravity", 1.0, W);
    const float dmin[4] = {-1, -1, 0, 0}, ddim[4] = {2, 2, 0, 0};
    as_vec4_or(root, "min", dmin, S.min, W);
    as_vec4_or(root, "dim", ddim, S.dim, W);
    S.render_h = S.h * (float)as_float_or(root, "particle_render_factor", 1.0, W);
    {
        float t[4];
        const float lp[4] = {-3, 3, -3, 0}, lc[4] = {1, 1, 1, 0};
        as_vec4_or(root, "light_pos", lp, t, W);
        std::memcpy(S.light_pos, t, 12);
        S.ambient = (float)as_float_or(root, "ambient_brightness", 1.0, W);
        S.diffuse = (float)as_float_or(root, "diffuse_brightness", 10.0, W);
        S.specular = (float)as_float_or(root, "specular_brightness", 0.0, W);
        as_vec4_or(root, "light_color", lc, t, W);
        std::memcpy(S.light_color, t, 12);
    }
    const float h = S.h, alpha = S.alpha;
    InterMap imap;

    const Value* bnds = root.get("boundaries");
    if (bnds && bnds->is_table()) {
        for (auto& kv : bnds->tab) {
            const std::string where = "[boundaries." + kv.first + "]";
            const Value& t = *kv.second;
            if (!t.is_table()) bad("%s must be a table", where.c_str());
            Object o;
            o.name = kv.first;
            o.packing = get_packing(t, h, where);
            const float friction = (float)as_float_or(t, "friction", 0.0, where);
            o.region = rp.from_mat(t, where);
            get_colors(t, o.colors, where);
            // Material::new_immobile (material.rs:113-120)
            o.mat.visc = friction;
            o.mat.start_den = 1.0f;
            o.mat.target_den = 1.0f;
            o.mat.immobile = 1;
            S.objects.push_back(o);
            get_interactions(o.name, t, (double)h, imap, where);
        }
    }

    const Value* objs = root.get("objects");
    if (objs && objs->is_table()) {
        for (auto& kv : objs->tab) {
            const std::string where = "[objects." + kv.first + "]";
            const Value& t = *kv.second;
            if (!t.is_table()) bad("%s must be a table", where.c_str());
            Object o;
            o.name = kv.first;
            ArpMaterial& m = o.mat;
            const double den = as_float_or(t, "density", 1.0, where);
            m.start_den = (float)as_float_or(t, "start_density", den, where);
            m.target_den = (float)as_float_or(t, "target_density", den, where);
            m.sound_speed = (float)as_float_or(t, "speed_of_sound", 0.0, where);
            m.visc = (float)as_float_or(t, "viscocity", as_float_or(t, "friction", 0.0, where), where);
            m.bulk_visc = (float)as_float_or(t, "bulk_viscocity", (double)(alpha * h * m.sound_speed), where);
            const Value* se = t.get("state");
            if (!se) se = t.get("state_equation");
            if (!se) {
                if (t.get("speed_of_sound"))
                    m.state_eq = (t.get("start_density") || t.get("target_density")) ? 2u : 3u;
                else
                    m.state_eq = 0;
            } else {
                if (!se->is_str()) bad("%s: the state equation must be a string", where.c_str());
                const std::string& s = se->str;
                if (s == "Tait" || s == "tait" || s == "Liquid" || s == "liquid") m.state_eq = 3;
                else if (s == "Ideal_Gas" || s == "ideal_gas" || s == "gas" || s == "Gas") m.state_eq = 2;
                else if (s == "Constant" || s == "constant") m.state_eq = 1;
                else bad("%s: invalid state equation: %s", where.c_str(), s.c_str());
            }
            const int64_t so = as_int_or(t, "strain_order", 2, where);
            m.strain_order = (int32_t)so;
            m.normal_stiffness = (float)as_float_or(t, "normal_stiffness", 0.0, where);
            m.shear_stiffness = (float)as_float_or(t, "shear_stiffness", 0.0, where);
            m.normal_damp = (float)as_float_or(t, "normal_dampening", 0.0, where);
            m.shear_damp = (float)as_float_or(t, "shear_dampening", 0.0, where);
            if (t.get("yield_strength") || t.get("relaxation_time")) {
                m.plastic = 1;
                m.yield_strength = (float)as_float_or(t, "yield_strength", 0.0, where);
                m.work_hardening = (float)as_float_or(t, "work_hardening", 0.0, where);
                m.work_hardening_exp = (float)as_float_or(t, "work_hardening_exp", 2.0, where);
                m.kinematic_hardening = (float)as_float_or(t, "kinematic_hardening", 0.0, where);
                m.thermal_softening = (float)as_float_or(t, "thermal_softening", 0.0, where);
                m.relaxation_time = (float)as_float_or(t, "relaxation_time", 1.0, where);
            }
            o.region = rp.from_mat(t, where);
            const float z4[4] = {0, 0, 0, 0};
            as_vec4_or(t, "velocity", z4, o.vel, where);
            o.region->bound().center(o.center);
            const Value* av = t.get("angular_velocity");
            if (av) {
                if (!av->is_array()) bad("%s: 'angular_velocity' must be an array", where.c_str());
                for (size_t i = 0; i < av->arr.size() && i < 6; i++) {
                    if (!av->arr[i]->is_float()) bad("%s: entries of 'angular_velocity' must be floats", where.c_str());
                    o.ang[i] = (float)av->arr[i]->f;
                }
            }
            get_colors(t, o.colors, where);
            o.packing = get_packing(t, h, where);
            o.moving = true;
            const Value* oc = t.get("on_click");
            if (!oc) {
                S.objects.push_back(o);
                get_interactions(o.name, t, (double)h, imap, where);
            } else {
                if (!oc->is_str()) bad("%s: 'on_click' must be a string", where.c_str());
                o.relative = oc->str == "Relative";
                S.on_click.push_back(o);
            }
        }
    }

    S.integrator = as_str_or(root, "integrator", "verlet", W);
    const size_t no = S.objects.size();
    S.inter_cfg.assign(no * no, ArpMatInteraction{});
    S.inter_present.assign(no * no, 0);
    for (size_t i = 0; i < no; i++)
        for (size_t j = 0; j < no; j++) {
            const ArpMatInteraction* a = find_inter(imap, S.objects[i].name, S.objects[j].name);
            if (!a) a = find_inter(imap, S.objects[j].name, S.objects[i].name);
            if (a) {
                S.inter_cfg[i * no + j] = *a;
                S.inter_present[i * no + j] = 1;
            }
        }
    const Value* vp = root.get("view_pos");
    S.view_pos = {0.0, 0.0};
    if (vp) {
        if (!vp->is_array()) bad("scene: 'view_pos' must be an array");
        S.view_pos.clear();
        for (auto& e : vp->arr) {
            if (!e->is_float()) bad("scene: entries of 'view_pos' must be floats");
            S.view_pos.push_back(e->f);
        }
    }
    S.view_scale = as_float_or(root, "view_scale", 1.0, W);
    S.view_angle = as_float_or(root, "view_angle", 0.0, W) * (kPi / 180.0);
}

// ---- lattice (material_region/mod.rs:56-155) ---------------------------------------------
constexpr double kMaxLattice = 4294967295.0;
constexpr size_t kBlock = (size_t)1 << 20;  // lattice points tested per batch

// The staggered lattice over the region's bounding box, in the reference's f32 order; the
// points are kept when the region holds them (tested in batches on `threads` threads; the
// order is kept).
void gen_object(const Object& o, float kernel_radius, uint32_t mat_id, unsigned threads,
                std::vector<ArpParticle>& out, ArpMaterial& mat, uint32_t& tested) {
    const float h = o.packing * kernel_radius;
    if (!(h > 0.0f) || !std::isfinite(h)) bad("object '%s': lattice spacing %g is not positive", o.name.c_str(), h);
    const Box bound = o.region->bound();
    for (int i = 0; i < 4; i++)
        if (!std::isfinite(bound.min[i]) || !std::isfinite(bound.dim[i]))
            bad("object '%s': region bound is not finite", o.name.c_str());
    {  // the lattice size, bounded before the walk
        double cnt = 1.0;
        for (int i = 0; i < 3; i++) {
            const double step = i == 0 ? h : i == 1 ? h / std::sqrt(2.0) : h / std::sqrt(3.0);
            if (bound.dim[i] > 0.0f) cnt *= std::floor(bound.dim[i] / step) + 2.0;
        }
        if (cnt > kMaxLattice) bad("object '%s': lattice of ~%.3g points exceeds 2^32", o.name.c_str(), cnt);
    }
    const float* c = o.center;
    const float* w = o.ang;
    const float* v = o.vel;
    std::vector<float> pts;
    pts.reserve(4 * kBlock);
    std::vector<uint8_t> keep;
    auto flush = [&]() {
        const size_t m = pts.size() / 4;
        if (!m) return;
        keep.assign(m, 0);
        auto work = [&](size_t b, size_t e) {
            for (size_t k = b; k < e; k++) keep[k] = o.region->contains(&pts[4 * k]) ? 1 : 0;
        };
        const unsigned nt = std::max(1u, std::min<unsigned>(threads, (unsigned)((m + 1023) / 1024)));
        if (nt == 1) {
            work(0, m);
        } else {
            std::vector<std::thread> th;
            const size_t chunk = (m + nt - 1) / nt;
            for (unsigned t = 0; t < nt; t++) {
                const size_t b = t * chunk, e = std::min(m, b + chunk);
                if (b < e) th.emplace_back(work, b, e);
            }
            for (auto& t : th) t.join();
        }
        for (size_t k = 0; k < m; k++) {
            if (!keep[k]) continue;
            const float* p = &pts[4 * k];
            ArpParticle q{};
            q.den = o.mat.start_den;
            q.mat = mat_id;
            q.solid_id = ARP_INVALID_INDEX;
            std::memcpy(q.pos, p, 16);
            if (o.moving) {  // the rigid velocity field of config.rs:405-418
                const float r[4] = {p[0] - c[0], p[1] - c[1], p[2] - c[2], p[3] - c[3]};
                q.vel[0] = w[1] * r[2] - w[2] * r[1] - w[3] * r[3] + v[0];
                q.vel[1] = w[2] * r[0] - w[0] * r[2] + w[4] * r[3] + v[1];
                q.vel[2] = w[0] * r[1] - w[1] * r[0] - w[5] * r[3] + v[2];
                q.vel[3] = w[3] * r[0] - w[4] * r[1] + w[5] * r[3] + v[3];
            }
            out.push_back(q);
        }
        pts.clear();
    };
    const float sq2 = std::sqrt(2.0f), sq3 = std::sqrt(3.0f), sq8 = std::sqrt(8.0f);
    float pos[4] = {bound.min[0], bound.min[1], bound.min[2], bound.min[3]};
    float offset2 = 0.0f;
    size_t num = 0;
    while (true) {
        float offset = 0.0f;
        pos[1] = bound.min[1] + offset2;
        while (true) {
            pos[0] = bound.min[0] + (offset + offset2);
            while (true) {
                if (pts.size() >= 4 * kBlock) flush();
                pts.insert(pts.end(), pos, pos + 4);
                num++;
                pos[0] += h;
                if (pos[0] - bound.min[0] > bound.dim[0] || bound.dim[0] == 0.0f) break;
            }
            pos[1] += h / sq2;
            offset = offset == 0.0f ? h / 2.0f : 0.0f;
            if (pos[1] - bound.min[1] > bound.dim[1] || bound.dim[1] == 0.0f) break;
        }
        pos[2] += h / sq3;
        offset2 = offset2 == 0.0f ? h / sq8 : 0.0f;
        if (pos[2] - bound.min[2] > bound.dim[2] || bound.dim[2] == 0.0f) break;
    }
    flush();
    double content = 1.0;
    for (int i = 0; i < 4; i++)
        if (bound.dim[i] > 0.0f) content = content * (double)bound.dim[i];
    const double box_mass = (double)o.mat.start_den * content;
    mat = o.mat;
    mat.mass = (float)(box_mass / (double)num);
    tested = (uint32_t)std::min<size_t>(num, 0xFFFFFFFFu);
}

void generate(arp_scene_file& S, unsigned threads) {
    S.particles.clear();
    S.boundary.clear();
    S.solids.clear();
    S.mats.assign(S.objects.size(), ArpMaterial{});
    S.first.assign(S.objects.size(), 0);
    S.count.assign(S.objects.size(), 0);
    S.tested.assign(S.objects.size(), 0);
    if (S.objects.size() > 0xFFFF) bad("too many objects");
    for (size_t i = 0; i < S.objects.size(); i++) {
        const Object& o = S.objects[i];
        std::vector<ArpParticle>& dst = o.mat.immobile ? S.boundary : S.particles;
        const size_t before = dst.size();
        gen_object(o, S.h, (uint32_t)i, threads, dst, S.mats[i], S.tested[i]);
        S.first[i] = (uint32_t)before;
        S.count[i] = (uint32_t)(dst.size() - before);
        if (dst.size() > 0xFFFFFFF0u) bad("scene exceeds 2^32 particles");
    }
    // init_solid_particles (particles.rs:488-507)
    for (size_t k = 0; k < S.particles.size(); k++) {
        ArpParticle& p = S.particles[k];
        const ArpMaterial& m = S.mats[p.mat];
        if (m.normal_stiffness != 0.0f || m.shear_stiffness != 0.0f) {
            p.solid_id = (uint32_t)S.solids.size();
            ArpSolidParticle s{};
            s.part_id = (uint32_t)k;
            std::memcpy(s.ref_pos, p.pos, 16);
            S.solids.push_back(s);
        }
    }
    if (S.solids.empty()) {
        ArpSolidParticle s{};
        s.part_id = ARP_INVALID_INDEX;
        S.solids.push_back(s);
    }
    // FluidSim::new's interaction matrix (soft_body/mod.rs:93-111)
    const uint32_t nm = (uint32_t)S.mats.size();
    S.inter.assign((size_t)nm * nm, ArpMatInteraction{});
    if (nm)
        arp_resolve_interactions(nm, S.mats.data(), S.inter_cfg.data(), S.inter_present.data(), S.h, S.inter.data());
    S.generated = true;
}

void copy_str(char* dst, size_t cap, const std::string& s) {
    const size_t n = std::min(cap - 1, s.size());
    std::memcpy(dst, s.data(), n);
    dst[n] = '\0';
}

int set_err(arp_scene_file* S, int rc, const std::string& m) {
    if (S) S->err = m;
    return rc;
}

}  // namespace

extern "C" {

int arp_scene_parse(const char* toml_text, size_t len, const char* base_dir, arp_scene_file** out, char* err,
                    size_t err_len) {
    auto report = [&](const std::string& m) {
        if (err && err_len) copy_str(err, err_len, m);
        return ARP_E_ARG;
    };
    if (!out) return report("arp_scene_parse: null output");
    *out = nullptr;
    if (!toml_text) return report("arp_scene_parse: null text");
    std::string perr;
    arp_toml::VPtr root = arp_toml::parse(toml_text, len, perr);
    if (!root) return report(perr);
    auto S = std::make_unique<arp_scene_file>();
    RegionParser rp;
    if (base_dir) rp.base_dir = base_dir;
    try {
        parse_config(*S, *root, rp);
    } catch (SceneError& e) {
        return report(e.msg);
    } catch (std::bad_alloc&) {
        return report("out of host memory");
    }
    if (err && err_len) err[0] = '\0';
    *out = S.release();
    return ARP_OK;
}

void arp_scene_free(arp_scene_file* f) { delete f; }

const char* arp_scene_error(const arp_scene_file* f) { return f ? f->err.c_str() : "null scene file"; }

int arp_scene_settings(const arp_scene_file* f, ArpSceneSettings* out) {
    if (!f || !out) return ARP_E_ARG;
    std::memset(out, 0, sizeof *out);
    copy_str(out->title, sizeof out->title, f->title);
    copy_str(out->integrator, sizeof out->integrator, f->integrator);
    out->kernel_radius = f->h;
    out->time_step = f->dt;
    out->artificial_viscosity = f->alpha;
    out->gravity = f->g;
    out->render_radius = f->render_h;
    out->subticks = f->subticks;
    std::memcpy(out->bounds.min, f->min, 16);
    std::memcpy(out->bounds.dim, f->dim, 16);
    out->n_objects = (uint32_t)f->objects.size();
    out->n_on_click = (uint32_t)f->on_click.size();
    out->lighting = f->lighting ? 1u : 0u;
    if (f->generated) {
        out->n_particles = (uint32_t)f->particles.size();
        out->n_boundary = (uint32_t)f->boundary.size();
        out->n_solids = (uint32_t)f->solids.size();
    }
    return ARP_OK;
}

int arp_scene_object(const arp_scene_file* f, uint32_t i, ArpSceneObject* out) {
    if (!f || !out) return ARP_E_ARG;
    const bool click = i >= f->objects.size();
    const size_t k = click ? i - f->objects.size() : i;
    if (click && k >= f->on_click.size()) return ARP_E_ARG;
    const Object& o = click ? f->on_click[k] : f->objects[k];
    std::memset(out, 0, sizeof *out);
    copy_str(out->name, sizeof out->name, o.name);
    std::memcpy(out->colors, o.colors, sizeof out->colors);
    out->packing = o.packing;
    out->immobile = o.mat.immobile;
    out->on_click = click ? (o.relative ? 2u : 1u) : 0u;
    const Box b = o.region->bound();
    std::memcpy(out->bound.min, b.min, 16);
    std::memcpy(out->bound.dim, b.dim, 16);
    out->material = o.mat;
    if (!click && f->generated) {
        out->material = f->mats[k];
        out->first = f->first[k];
        out->count = f->count[k];
        out->lattice_points = f->tested[k];
    }
    return ARP_OK;
}

int arp_scene_generate(arp_scene_file* f, uint32_t threads) {
    if (!f) return ARP_E_ARG;
    unsigned nt = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    nt = std::min(nt, 64u);
    try {
        generate(*f, nt);
    } catch (SceneError& e) {
        f->generated = false;
        return set_err(f, ARP_E_LIMIT, e.msg);
    } catch (std::bad_alloc&) {
        f->generated = false;
        return set_err(f, ARP_E_NOMEM, "out of host memory generating the scene");
    } catch (std::system_error&) {
        f->generated = false;
        return set_err(f, ARP_E_ARG, "could not start worker threads");
    }
    return ARP_OK;
}

int arp_scene_arrays(const arp_scene_file* f, ArpParticle* particles, uint32_t n, ArpSolidParticle* solids,
                     uint32_t ns, ArpParticle* boundary, uint32_t nb, ArpMaterial* mats, uint32_t nm,
                     ArpMatInteraction* inter) {
    if (!f) return ARP_E_ARG;
    if (!f->generated) return ARP_E_STATE;
    if ((particles && n != f->particles.size()) || (solids && ns != f->solids.size()) ||
        (boundary && nb != f->boundary.size()) || (mats && nm != f->mats.size()))
        return ARP_E_ARG;
    if (particles && n) std::memcpy(particles, f->particles.data(), sizeof(ArpParticle) * n);
    if (solids && ns) std::memcpy(solids, f->solids.data(), sizeof(ArpSolidParticle) * ns);
    if (boundary && nb) std::memcpy(boundary, f->boundary.data(), sizeof(ArpParticle) * nb);
    if (mats && nm) std::memcpy(mats, f->mats.data(), sizeof(ArpMaterial) * nm);
    if (inter && !f->inter.empty()) std::memcpy(inter, f->inter.data(), sizeof(ArpMatInteraction) * f->inter.size());
    return ARP_OK;
}

int arp_scene_contains(const arp_scene_file* f, uint32_t object, const float* points, uint32_t n, uint8_t* out) {
    if (!f || (!points && n) || (!out && n)) return ARP_E_ARG;
    const size_t no = f->objects.size();
    if (object >= no + f->on_click.size()) return ARP_E_ARG;
    const Object& o = object < no ? f->objects[object] : f->on_click[object - no];
    for (uint32_t k = 0; k < n; k++) out[k] = o.region->contains(points + 4 * (size_t)k) ? 1 : 0;
    return ARP_OK;
}

int arp_scene_apply(arp_ctx* ctx, const arp_scene_file* f) {
    if (!ctx || !f) return ARP_E_ARG;
    if (!f->generated) return ARP_E_STATE;
    ArpTableau tab;
    if (arp_tableau_by_name(f->integrator.c_str(), &tab) != ARP_OK) return ARP_E_ARG;
    ArpAABB b;
    std::memcpy(b.min, f->min, 16);
    std::memcpy(b.dim, f->dim, 16);
    int rc = arp_scene_upload(ctx, f->particles.data(), (uint32_t)f->particles.size(), f->solids.data(),
                              (uint32_t)f->solids.size(), f->boundary.data(), (uint32_t)f->boundary.size(),
                              f->mats.data(), (uint32_t)f->mats.size(), f->inter.data(), &b, f->h, f->g,
                              f->alpha);
    if (rc != ARP_OK) return rc;
    return arp_set_integrator(ctx, &tab, f->dt, f->subticks);
}

int arp_scene_spawn(arp_ctx* ctx, const arp_scene_file* f, uint32_t object, const float* offset,
                    uint32_t* material_added) {
    if (!ctx || !f) return ARP_E_ARG;
    const size_t no = f->objects.size();
    if (object >= no + f->on_click.size()) return ARP_E_ARG;
    const Object& o = object < no ? f->objects[object] : f->on_click[object - no];
    const float h = arp_kernel_radius(ctx);
    if (!(h > 0.0f)) return ARP_E_STATE;
    std::vector<ArpParticle> p;
    ArpMaterial mat{};
    uint32_t tested = 0;
    try {
        gen_object(o, h, 0, std::max(1u, std::min(16u, std::thread::hardware_concurrency())), p, mat, tested);
    } catch (SceneError&) {
        return ARP_E_LIMIT;
    } catch (std::bad_alloc&) {
        return ARP_E_NOMEM;
    } catch (std::system_error&) {
        return ARP_E_ARG;
    }
    if (offset)
        for (ArpParticle& q : p)
            for (int i = 0; i < 4; i++) q.pos[i] += offset[i];
    return arp_add_particles(ctx, &mat, p.data(), (uint32_t)p.size(), material_added);
}

int arp_toml_to_json(const char* text, size_t len, char* out, size_t cap, size_t* needed) {
    std::string err;
    arp_toml::VPtr root = arp_toml::parse(text ? text : "", text ? len : 0, err);
    const std::string s = root ? arp_toml::to_json(*root) : err;
    if (needed) *needed = s.size() + 1;
    if (out && cap) copy_str(out, cap, s);
    return root ? ARP_OK : ARP_E_ARG;
}

}  // extern "C"
