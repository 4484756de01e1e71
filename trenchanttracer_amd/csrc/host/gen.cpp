// Procedural scene and ray generators for the workloads (the reference loads
// OBJ/JSON assets that are not shipped; these replace them with seeded geometry).
// Geometry is computed in double precision and rounded once to float.
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "../common/philox.h"
#include "tt_host.h"

namespace tt {
namespace {

struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 dnorm(D3 a) {
    double l = std::sqrt(ddot(a, a));
    return l > 0 ? a * (1.0 / l) : D3{0, 0, 1};
}
inline V3 tof(D3 a) { return {(float)a.x, (float)a.y, (float)a.z}; }
inline D3 tod(V3 a) { return {a.x, a.y, a.z}; }

const double kPi = 3.14159265358979323846;

struct Rng {
    uint32_t k0, k1;
    explicit Rng(uint64_t seed) : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)) {}
    u32x4 draw(uint32_t a, uint32_t b, uint32_t c = 0, uint32_t d = 0) const {
        return philox4x32_10({a, b, c, d}, k0, k1);
    }
};

// Orthonormal frame around unit n (Duff et al. 2017, "Building an orthonormal basis, revisited").
void frame(D3 n, D3 &t, D3 &b) {
    double s = n.z >= 0 ? 1.0 : -1.0;
    double a = -1.0 / (s + n.z);
    double bb = n.x * n.y * a;
    t = {1.0 + s * n.x * n.x * a, s * bb, -s * n.x};
    b = {bb, s + n.y * n.y * a, -n.y};
}

// Value noise on the integer lattice, smoothstep-interpolated, in [-1, 1].
double lattice(const Rng &r, int64_t x, int64_t y, int64_t z, uint32_t oct) {
    u32x4 h = r.draw((uint32_t)x, (uint32_t)y, (uint32_t)z, 0x0C7A0000u + oct);
    return u01d(h.x, h.y) * 2.0 - 1.0;
}
double value_noise(const Rng &r, D3 p, uint32_t oct) {
    double fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double tx = p.x - fx, ty = p.y - fy, tz = p.z - fz;
    tx = tx * tx * (3 - 2 * tx);
    ty = ty * ty * (3 - 2 * ty);
    tz = tz * tz * (3 - 2 * tz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(r, ix + a, iy + b, iz + d, oct);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
    double x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
    double x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty;
    double y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}
double fbm(const Rng &r, D3 p, int octaves) {
    double sum = 0, amp = 0.5, freq = 2.0;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * value_noise(r, p * freq, (uint32_t)o);
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum;
}

}  // namespace

void face_normal_fill(Mesh &m) {
    // One face normal per triangle: normalize((v1 - v0) x (v2 - v0)).
    const int nt = m.ntris();
    m.normals.resize(nt);
    m.tri_n.resize(3 * (size_t)nt);
    for (int i = 0; i < nt; ++i) {
        D3 v0 = tod(m.verts[m.tri_v[3 * i]]), v1 = tod(m.verts[m.tri_v[3 * i + 1]]), v2 = tod(m.verts[m.tri_v[3 * i + 2]]);
        m.normals[i] = tof(dnorm(dcross(v1 - v0, v2 - v0)));
        m.tri_n[3 * i] = m.tri_n[3 * i + 1] = m.tri_n[3 * i + 2] = i;
    }
}

void gen_soup(int32_t ntris, uint64_t seed, float mu, float sigma, Mesh &m) {
    Rng r(seed);
    m = Mesh();
    m.verts.resize(3 * (size_t)ntris);
    m.tri_v.resize(3 * (size_t)ntris);
    m.tri_mat.assign(ntris, 0);
    for (int32_t i = 0; i < ntris; ++i) {
        u32x4 a = r.draw((uint32_t)i, 0);
        u32x4 b = r.draw((uint32_t)i, 1);
        u32x4 c = r.draw((uint32_t)i, 2);
        D3 centre = {u01d(a.x, a.y), u01d(a.z, a.w), u01d(b.x, b.y)};
        // side ~ LogNormal(mu, sigma) by Box-Muller
        double u1 = 1.0 - u01d(b.z, b.w), u2 = u01d(c.x, c.y);
        double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
        double side = std::exp((double)mu + (double)sigma * g);
        // uniform rotation (Shoemake 1992)
        u32x4 d = r.draw((uint32_t)i, 3);
        double q1 = u01d(c.z, c.w), q2 = u01d(d.x, d.y), q3 = u01d(d.z, d.w);
        double s1 = std::sqrt(1 - q1), s2 = std::sqrt(q1);
        double qx = s1 * std::sin(2 * kPi * q2), qy = s1 * std::cos(2 * kPi * q2);
        double qz = s2 * std::sin(2 * kPi * q3), qw = s2 * std::cos(2 * kPi * q3);
        double R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                          {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                          {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
        double circ = side / std::sqrt(3.0);
        for (int k = 0; k < 3; ++k) {
            double ang = kPi / 2 + k * 2 * kPi / 3;
            D3 l = {circ * std::cos(ang), circ * std::sin(ang), 0};
            D3 w = {R[0][0] * l.x + R[0][1] * l.y + R[0][2] * l.z, R[1][0] * l.x + R[1][1] * l.y + R[1][2] * l.z,
                    R[2][0] * l.x + R[2][1] * l.y + R[2][2] * l.z};
            m.verts[3 * (size_t)i + k] = tof(centre + w);
            m.tri_v[3 * (size_t)i + k] = 3 * i + k;
        }
    }
    face_normal_fill(m);
}

void gen_icosphere(int32_t subdiv, float radius, const float c[3], float amp, int32_t octaves, uint64_t seed,
                   int32_t mat, Mesh &m) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<D3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto &p : v) p = dnorm(p);
    std::vector<int32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                              11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                              3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int s = 0; s < subdiv; ++s) {
        std::unordered_map<uint64_t, int32_t> mid;
        mid.reserve(f.size());
        auto midpoint = [&](int32_t a, int32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | (uint32_t)b) : ((uint64_t)b << 32 | (uint32_t)a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            int32_t id = (int32_t)v.size();
            v.push_back(dnorm((v[a] + v[b]) * 0.5));
            mid.emplace(key, id);
            return id;
        };
        std::vector<int32_t> nf;
        nf.reserve(f.size() * 4);
        for (size_t i = 0; i < f.size(); i += 3) {
            int32_t a = f[i], b = f[i + 1], cc = f[i + 2];
            int32_t ab = midpoint(a, b), bc = midpoint(b, cc), ca = midpoint(cc, a);
            int32_t tri[12] = {a, ab, ca, b, bc, ab, cc, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), tri, tri + 12);
        }
        f.swap(nf);
    }
    Rng r(seed);
    std::vector<D3> p(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        double rr = radius * (1.0 + (amp != 0.0f ? (double)amp * fbm(r, v[i], octaves) : 0.0));
        p[i] = D3{c[0], c[1], c[2]} + v[i] * rr;
    }
    // Ensure outward CCW winding: (p1 - p0) x (p2 - p0) points away from the centre.
    {
        D3 e = dcross(p[f[1]] - p[f[0]], p[f[2]] - p[f[0]]);
        D3 ctr = {c[0], c[1], c[2]};
        if (ddot(e, p[f[0]] - ctr) < 0)
            for (size_t i = 0; i < f.size(); i += 3) std::swap(f[i + 1], f[i + 2]);
    }
    std::vector<D3> acc(v.size(), D3{0, 0, 0});
    for (size_t i = 0; i < f.size(); i += 3) {
        D3 n = dcross(p[f[i + 1]] - p[f[i]], p[f[i + 2]] - p[f[i]]);  // area weighted
        for (int k = 0; k < 3; ++k) acc[f[i + k]] = acc[f[i + k]] + n;
    }
    m = Mesh();
    m.verts.resize(v.size());
    m.normals.resize(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        m.verts[i] = tof(p[i]);
        m.normals[i] = tof(dnorm(acc[i]));
    }
    m.tri_v = f;
    m.tri_n = f;
    m.tri_mat.assign(f.size() / 3, mat);
}

void gen_quad(const float p0[3], const float p1[3], const float p2[3], int32_t mat, Mesh &m) {
    D3 a = {p0[0], p0[1], p0[2]}, b = {p1[0], p1[1], p1[2]}, c = {p2[0], p2[1], p2[2]};
    D3 d = b + c - a;
    m = Mesh();
    m.verts = {tof(a), tof(b), tof(c), tof(d)};
    m.tri_v = {0, 1, 2, 1, 3, 2};
    m.tri_mat = {mat, mat};
    D3 n = dnorm(dcross(b - a, c - a));
    m.normals = {tof(n)};
    m.tri_n = {0, 0, 0, 0, 0, 0};
}

void assign_materials(Mesh &m, int32_t k, const float *fr, const int32_t *mats, uint64_t seed) {
    Rng r(seed);
    for (int32_t i = 0; i < m.ntris(); ++i) {
        u32x4 h = r.draw((uint32_t)i, 0x3A7u);
        double u = u01d(h.x, h.y), acc = 0;
        int32_t pick = mats[k - 1];
        for (int j = 0; j < k; ++j) {
            acc += fr[j];
            if (u < acc) {
                pick = mats[j];
                break;
            }
        }
        m.tri_mat[i] = pick;
    }
}

void rays_primary(int32_t w, int32_t h, int32_t spp, const float eye[3], const float target[3], float fovx_deg,
                  float tmin, float tmax, uint64_t seed, tt_ray *out) {
    Rng r(seed);
    D3 e = {eye[0], eye[1], eye[2]};
    D3 view = dnorm(D3{target[0], target[1], target[2]} - e);
    D3 right = dnorm(dcross(view, D3{0, 1, 0}));
    D3 up = dcross(right, view);
    double tx = std::tan((double)fovx_deg * kPi / 360.0);
    double ty = tx * (double)h / (double)w;
    int64_t idx = 0;
    const int T = 8;
    for (int ty0 = 0; ty0 < h; ty0 += T)
        for (int tx0 = 0; tx0 < w; tx0 += T)
            for (int s = 0; s < spp; ++s)
                for (int yy = ty0; yy < std::min(h, ty0 + T); ++yy)
                    for (int xx = tx0; xx < std::min(w, tx0 + T); ++xx) {
                        u32x4 j = r.draw((uint32_t)(yy * w + xx), (uint32_t)s, 0x9A11u);
                        double sx = (xx + u01d(j.x, j.y)) / w * 2.0 - 1.0;
                        double sy = 1.0 - (yy + u01d(j.z, j.w)) / h * 2.0;
                        D3 d = dnorm(view + right * (sx * tx) + up * (sy * ty));
                        tt_ray &ry = out[idx++];
                        ry = {(float)e.x, (float)e.y, (float)e.z, tmin, (float)d.x, (float)d.y, (float)d.z, tmax};
                    }
}

void rays_incoherent(const Mesh &m, int64_t n, float tmin, float tmax, uint64_t seed, tt_ray *out) {
    Rng r(seed);
    const int32_t nt = m.ntris();
    for (int64_t i = 0; i < n; ++i) {
        u32x4 a = r.draw((uint32_t)i, (uint32_t)(i >> 32), 0x1C0u);
        u32x4 b = r.draw((uint32_t)i, (uint32_t)(i >> 32), 0x1C1u);
        int32_t tri = (int32_t)(u01d(a.x, a.y) * nt);
        if (tri >= nt) tri = nt - 1;
        D3 v0 = tod(m.verts[m.tri_v[3 * (size_t)tri]]), v1 = tod(m.verts[m.tri_v[3 * (size_t)tri + 1]]),
           v2 = tod(m.verts[m.tri_v[3 * (size_t)tri + 2]]);
        double r1 = std::sqrt(u01d(a.z, a.w)), r2 = u01d(b.x, b.y);
        D3 p = v0 * (1 - r1) + v1 * (r1 * (1 - r2)) + v2 * (r1 * r2);
        D3 nrm = dnorm(dcross(v1 - v0, v2 - v0));
        D3 tt, bb;
        frame(nrm, tt, bb);
        double u3 = u01d(b.z, b.w);
        u32x4 c = r.draw((uint32_t)i, (uint32_t)(i >> 32), 0x1C2u);
        double phi = 2 * kPi * u01d(c.x, c.y);
        double sr = std::sqrt(u3);
        D3 d = dnorm(tt * (sr * std::cos(phi)) + bb * (sr * std::sin(phi)) + nrm * std::sqrt(1 - u3));
        out[i] = {(float)p.x, (float)p.y, (float)p.z, tmin, (float)d.x, (float)d.y, (float)d.z, tmax};
    }
}

}  // namespace tt
