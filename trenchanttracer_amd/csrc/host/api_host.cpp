// C ABI, host half: meshes, generators, SBVH build / export / cache file.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <unordered_map>

#include "tt_host.h"

namespace tt {

void gen_soup(int32_t, uint64_t, float, float, Mesh &);
void gen_icosphere(int32_t, float, const float *, float, int32_t, uint64_t, int32_t, Mesh &);
void gen_quad(const float *, const float *, const float *, int32_t, Mesh &);
void assign_materials(Mesh &, int32_t, const float *, const int32_t *, uint64_t);
void rays_primary(int32_t, int32_t, int32_t, const float *, const float *, float, float, float, uint64_t, tt_ray *);
void rays_incoherent(const Mesh &, int64_t, float, float, uint64_t, tt_ray *);

namespace {
thread_local std::string g_err;
}

void set_error(const std::string &msg) { g_err = msg; }
tt_status fail(tt_status code, const std::string &msg) {
    g_err = msg;
    return code;
}

}  // namespace tt

using namespace tt;

#define TT_GUARD_BEGIN try {
#define TT_GUARD_END                                                    \
    }                                                                   \
    catch (const std::bad_alloc &) {                                    \
        return fail(TT_ERR_NOMEM, "host allocation failed");            \
    }                                                                   \
    catch (const std::exception &e) {                                   \
        return fail(TT_ERR_INVALID, std::string("exception: ") + e.what()); \
    }

static_assert(sizeof(tt_ray) == 32, "tt_ray must be 32 bytes");
static_assert(sizeof(tt_hit) == 16, "tt_hit must be 16 bytes");
static_assert(sizeof(tt_material) == 48, "tt_material must be 48 bytes");
static_assert(sizeof(tt_geometry) == 48, "tt_geometry must be 48 bytes");
static_assert(sizeof(tt_trace_cfg) == 24, "tt_trace_cfg must be 24 bytes");
static_assert(sizeof(tt_camera) == 60, "tt_camera must be 60 bytes");
static_assert(sizeof(tt_render_params) == 76, "tt_render_params must be 76 bytes");
static_assert(sizeof(tt_render_stats) == 40, "tt_render_stats must be 40 bytes");
static_assert(sizeof(tt_guide_params) == 36, "tt_guide_params must be 36 bytes");
static_assert(sizeof(tt_guide_stats) == 32, "tt_guide_stats must be 32 bytes");

namespace tt {

// The traversal kernel follows child indices from node 0 with a stack of 72 entries
// and walks a leaf's triangle records until one carries the end flag, so a layout
// from a file must be a tree (each node reached once: the walk ends) no deeper than
// 64 inner nodes, with every leaf starting inside the triangle array and the last
// record flagged (a leaf walk then stops inside the array).
tt_status check_device_layout(const Flat &f) {
    if (f.dev_nodes.size() % 16 != 0 || f.dev_tris.size() % 12 != 0)
        return fail(TT_ERR_IO, "bvh: device layout has a partial node or triangle record");
    const int64_t dn = (int64_t)(f.dev_nodes.size() / 16), dt = (int64_t)(f.dev_tris.size() / 12);
    if (dn == 0) return TT_OK;
    if (dt > 0) {
        int32_t last;
        std::memcpy(&last, &f.dev_tris[(size_t)(dt - 1) * 12 + 11], 4);
        if (last == 0) return fail(TT_ERR_IO, "bvh: the last triangle record does not end a leaf");
    }
    std::vector<uint8_t> seen((size_t)dn, 0);
    std::vector<std::pair<int64_t, int32_t>> stack{{0, 0}};
    seen[0] = 1;
    while (!stack.empty()) {
        const int64_t i = stack.back().first;
        const int32_t depth = stack.back().second;
        stack.pop_back();
        for (int c = 0; c < 2; ++c) {
            const int32_t ch = f.dev_nodes[(size_t)i * 16 + 12 + c];
            if (ch >= 0) {
                if (ch >= dn || seen[(size_t)ch]) return fail(TT_ERR_IO, "bvh: node graph is not a tree");
                if (depth + 1 > 64) return fail(TT_ERR_IO, "bvh: tree deeper than 64 levels");
                seen[(size_t)ch] = 1;
                stack.push_back({ch, depth + 1});
            } else if (~(int64_t)ch >= dt) {
                return fail(TT_ERR_IO, "bvh: leaf starts past the triangle array");
            }
        }
    }
    return TT_OK;
}

tt_status check_tri_ids(const Flat &f, int64_t ntris) {
    for (size_t q = 3; q < f.dev_tris.size(); q += 12) {
        int32_t id;
        std::memcpy(&id, &f.dev_tris[q], 4);
        if (id < -1 || id >= ntris) return fail(TT_ERR_INVALID, "bvh: triangle id outside the mesh");
    }
    return TT_OK;
}

}  // namespace tt

// Copies a host array out to a caller buffer (nothing for an empty array, whose
// data() may be null).
template <class V>
static void copy_out(void *dst, const V &v, size_t elem_bytes) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * elem_bytes);
}

extern "C" {

const char *tt_last_error(void) { return tt::g_err.c_str(); }
int tt_abi_version(void) { return 5; }   // 2: refill fields in tt_trace_cfg / tt_render_params; 3: max_bounce; 4: reorder; 5: tile, first_loss, pass_growth, learn_alpha

tt_status tt_mesh_create(const float *verts, int32_t nverts, const int32_t *tris, int32_t ntris,
                         const float *normals, int32_t nnormals, const int32_t *tri_normals,
                         const int32_t *tri_mats, tt_mesh **out) {
    TT_GUARD_BEGIN
    if (!out || nverts < 0 || ntris < 0 || (nverts > 0 && !verts) || (ntris > 0 && !tris))
        return fail(TT_ERR_INVALID, "tt_mesh_create: bad arguments");
    if (normals && (nnormals < 0 || (ntris > 0 && !tri_normals)))
        return fail(TT_ERR_INVALID, "tt_mesh_create: normals given without tri_normals");
    auto m = std::make_unique<tt_mesh>();
    Mesh &M = m->m;
    M.verts.resize(nverts);
    for (int32_t i = 0; i < nverts; ++i) M.verts[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    M.tri_v.assign(tris, tris + 3 * (size_t)ntris);
    for (int32_t i = 0; i < 3 * ntris; ++i)
        if (M.tri_v[i] < 0 || M.tri_v[i] >= nverts) return fail(TT_ERR_INVALID, "tt_mesh_create: vertex index out of range");
    M.tri_mat.assign(ntris, 0);
    if (tri_mats) M.tri_mat.assign(tri_mats, tri_mats + ntris);
    if (normals) {
        M.normals.resize(nnormals);
        for (int32_t i = 0; i < nnormals; ++i) M.normals[i] = {normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]};
        M.tri_n.assign(tri_normals, tri_normals + 3 * (size_t)ntris);
        for (int32_t i = 0; i < 3 * ntris; ++i)
            if (M.tri_n[i] < 0 || M.tri_n[i] >= nnormals) return fail(TT_ERR_INVALID, "tt_mesh_create: normal index out of range");
    } else {
        face_normal_fill(M);
    }
    *out = m.release();
    return TT_OK;
    TT_GUARD_END
}

void tt_mesh_destroy(tt_mesh *m) { delete m; }

tt_status tt_mesh_info(const tt_mesh *m, int32_t *nverts, int32_t *ntris, int32_t *nnormals) {
    if (!m) return fail(TT_ERR_INVALID, "tt_mesh_info: null mesh");
    if (nverts) *nverts = (int32_t)m->m.verts.size();
    if (ntris) *ntris = m->m.ntris();
    if (nnormals) *nnormals = (int32_t)m->m.normals.size();
    return TT_OK;
}

tt_status tt_mesh_get(const tt_mesh *m, float *verts, int32_t *tris, float *normals, int32_t *tri_normals,
                      int32_t *tri_mats) {
    if (!m) return fail(TT_ERR_INVALID, "tt_mesh_get: null mesh");
    const Mesh &M = m->m;
    copy_out(verts, M.verts, 12);
    copy_out(tris, M.tri_v, 4);
    copy_out(normals, M.normals, 12);
    copy_out(tri_normals, M.tri_n, 4);
    copy_out(tri_mats, M.tri_mat, 4);
    return TT_OK;
}

tt_status tt_mesh_append(tt_mesh *dst, const tt_mesh *src, int32_t mat_offset) {
    TT_GUARD_BEGIN
    if (!dst || !src) return fail(TT_ERR_INVALID, "tt_mesh_append: null mesh");
    Mesh &D = dst->m;
    const Mesh &S = src->m;
    int32_t vb = (int32_t)D.verts.size(), nb = (int32_t)D.normals.size();
    D.verts.insert(D.verts.end(), S.verts.begin(), S.verts.end());
    D.normals.insert(D.normals.end(), S.normals.begin(), S.normals.end());
    for (int32_t x : S.tri_v) D.tri_v.push_back(x + vb);
    for (int32_t x : S.tri_n) D.tri_n.push_back(x + nb);
    for (int32_t x : S.tri_mat) D.tri_mat.push_back(x + mat_offset);
    return TT_OK;
    TT_GUARD_END
}

tt_status tt_mesh_gen_soup(int32_t ntris, uint64_t seed, float mu, float sigma, tt_mesh **out) {
    TT_GUARD_BEGIN
    if (!out || ntris < 0 || ntris > (1 << 28)) return fail(TT_ERR_INVALID, "tt_mesh_gen_soup: bad arguments");
    auto m = std::make_unique<tt_mesh>();
    gen_soup(ntris, seed, mu, sigma, m->m);
    *out = m.release();
    return TT_OK;
    TT_GUARD_END
}

tt_status tt_mesh_gen_icosphere(int32_t subdiv, float radius, const float c[3], float amp, int32_t octaves,
                                uint64_t seed, int32_t mat, tt_mesh **out) {
    TT_GUARD_BEGIN
    if (!out || !c || subdiv < 0 || subdiv > 10 || octaves < 0 || octaves > 16)
        return fail(TT_ERR_INVALID, "tt_mesh_gen_icosphere: bad arguments");
    auto m = std::make_unique<tt_mesh>();
    gen_icosphere(subdiv, radius, c, amp, octaves, seed, mat, m->m);
    *out = m.release();
    return TT_OK;
    TT_GUARD_END
}

tt_status tt_mesh_gen_quad(const float p0[3], const float p1[3], const float p2[3], int32_t mat, tt_mesh **out) {
    TT_GUARD_BEGIN
    if (!out || !p0 || !p1 || !p2) return fail(TT_ERR_INVALID, "tt_mesh_gen_quad: bad arguments");
    auto m = std::make_unique<tt_mesh>();
    gen_quad(p0, p1, p2, mat, m->m);
    *out = m.release();
    return TT_OK;
    TT_GUARD_END
}

tt_status tt_mesh_assign_materials(tt_mesh *m, int32_t k, const float *fractions, const int32_t *mats,
                                   uint64_t seed) {
    if (!m || k <= 0 || !fractions || !mats) return fail(TT_ERR_INVALID, "tt_mesh_assign_materials: bad arguments");
    assign_materials(m->m, k, fractions, mats, seed);
    return TT_OK;
}

tt_status tt_rays_primary(int32_t width, int32_t height, int32_t spp, const float eye[3], const float target[3],
                          float fovx_deg, float tmin, float tmax, uint64_t seed, tt_ray *out) {
    if (width <= 0 || height <= 0 || spp <= 0 || !eye || !target || !out || !(fovx_deg > 0 && fovx_deg < 180))
        return fail(TT_ERR_INVALID, "tt_rays_primary: bad arguments");
    rays_primary(width, height, spp, eye, target, fovx_deg, tmin, tmax, seed, out);
    return TT_OK;
}

tt_status tt_rays_incoherent(const tt_mesh *m, int64_t n, float tmin, float tmax, uint64_t seed, tt_ray *out) {
    if (!m || n < 0 || (n > 0 && !out) || (n > 0 && m->m.ntris() == 0))
        return fail(TT_ERR_INVALID, "tt_rays_incoherent: bad arguments");
    rays_incoherent(m->m, n, tmin, tmax, seed, out);
    return TT_OK;
}

tt_status tt_bvh_build(const tt_mesh *m, float split_alpha, int32_t nthreads, tt_bvh **out) {
    TT_GUARD_BEGIN
    if (!m || !out || !(split_alpha >= 0.0f)) return fail(TT_ERR_INVALID, "tt_bvh_build: bad arguments");
    auto b = std::make_unique<tt_bvh>();
    build_sbvh(m->m, split_alpha, nthreads, b->b);
    flatten(m->m, b->b);
    *out = b.release();
    return TT_OK;
    TT_GUARD_END
}

void tt_bvh_destroy(tt_bvh *b) { delete b; }

tt_status tt_bvh_get_stats(const tt_bvh *b, tt_bvh_stats *out) {
    if (!b || !out) return fail(TT_ERR_INVALID, "tt_bvh_get_stats: bad arguments");
    *out = b->b.stats;
    return TT_OK;
}

tt_status tt_bvh_export_compact(const tt_bvh *b, tt_compact_sizes *sizes, int32_t *nodes, float *vertices,
                                float *normals, int32_t *tri_indices, int32_t *mat_indices) {
    if (!b || !sizes) return fail(TT_ERR_INVALID, "tt_bvh_export_compact: bad arguments");
    const Flat &f = b->b.flat;
    sizes->nodes = (uint32_t)(f.ref_nodes.size() / 4);
    sizes->vertices = (uint32_t)(f.ref_vertices.size() / 4);
    sizes->normals = (uint32_t)(f.ref_normals.size() / 4);
    sizes->tri_indices = (uint32_t)f.ref_tri_idx.size();
    sizes->mat_indices = (uint32_t)f.ref_mat_idx.size();
    sizes->leaf_count = f.leaf_count;
    sizes->tri_count = f.tri_count;
    copy_out(nodes, f.ref_nodes, 4);
    copy_out(vertices, f.ref_vertices, 4);
    copy_out(normals, f.ref_normals, 4);
    copy_out(tri_indices, f.ref_tri_idx, 4);
    copy_out(mat_indices, f.ref_mat_idx, 4);
    return TT_OK;
}

tt_status tt_bvh_export_device_layout(const tt_bvh *b, int64_t *num_nodes, int64_t *num_tris, int32_t *nodes,
                                      float *tris) {
    if (!b) return fail(TT_ERR_INVALID, "tt_bvh_export_device_layout: null bvh");
    const Flat &f = b->b.flat;
    if (num_nodes) *num_nodes = (int64_t)(f.dev_nodes.size() / 16);
    if (num_tris) *num_tris = (int64_t)(f.dev_tris.size() / 12);
    copy_out(nodes, f.dev_nodes, 4);
    copy_out(tris, f.dev_tris, 4);
    return TT_OK;
}

// ---------------------------------------------------------------- cache file
// Layout: the reference bvh.cache (7 x u32 header {nodesSize, triCount,
// leafNodeCount, verticesSize, normalsSize, triIndicesSize, matIndicesSize}, then
// the five arrays, BVHCompact.cpp:319-375), then a trailer
// {"TTDL", u32 version, i64 dev_nodes, i64 dev_tris, stats, node words, tri floats}.
// A file without the trailer (a reference cache) is imported by converting the
// compact arrays to the device layout.

static bool wr(FILE *f, const void *p, size_t bytes) { return bytes == 0 || fwrite(p, 1, bytes, f) == bytes; }
static bool rd(FILE *f, void *p, size_t bytes) { return bytes == 0 || fread(p, 1, bytes, f) == bytes; }

tt_status tt_bvh_save(const tt_bvh *b, const char *path) {
    TT_GUARD_BEGIN
    if (!b || !path) return fail(TT_ERR_INVALID, "tt_bvh_save: bad arguments");
    const Flat &f = b->b.flat;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(TT_ERR_IO, std::string("tt_bvh_save: cannot open ") + path);
    uint32_t hdr[7] = {(uint32_t)(f.ref_nodes.size() / 4), f.tri_count, f.leaf_count,
                       (uint32_t)(f.ref_vertices.size() / 4), (uint32_t)(f.ref_normals.size() / 4),
                       (uint32_t)f.ref_tri_idx.size(), (uint32_t)f.ref_mat_idx.size()};
    int64_t dn = (int64_t)(f.dev_nodes.size() / 16), dt = (int64_t)(f.dev_tris.size() / 12);
    uint32_t ver = 1;
    bool ok = wr(fp, hdr, sizeof(hdr)) && wr(fp, f.ref_nodes.data(), f.ref_nodes.size() * 4) &&
              wr(fp, f.ref_vertices.data(), f.ref_vertices.size() * 4) &&
              wr(fp, f.ref_normals.data(), f.ref_normals.size() * 4) &&
              wr(fp, f.ref_tri_idx.data(), f.ref_tri_idx.size() * 4) &&
              wr(fp, f.ref_mat_idx.data(), f.ref_mat_idx.size() * 4) && wr(fp, "TTDL", 4) && wr(fp, &ver, 4) &&
              wr(fp, &dn, 8) && wr(fp, &dt, 8) && wr(fp, &b->b.stats, sizeof(tt_bvh_stats)) &&
              wr(fp, f.dev_nodes.data(), f.dev_nodes.size() * 4) && wr(fp, f.dev_tris.data(), f.dev_tris.size() * 4);
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) return fail(TT_ERR_IO, std::string("tt_bvh_save: write failed for ") + path);
    return TT_OK;
    TT_GUARD_END
}

// Rebuilds the device layout from reference compact arrays: leaves are laid out
// in increasing vertex offset (the order createCompact emitted them), each
// leaf's vertex run walked up to its terminator.
static tt_status compact_to_device(Flat &f) {
    const size_t nn = f.ref_nodes.size() / 16;
    const size_t nvert = f.ref_vertices.size() / 4;
    f.dev_nodes.assign(f.ref_nodes.begin(), f.ref_nodes.end());
    f.dev_tris.clear();
    std::vector<int32_t> offs;
    for (size_t k = 0; k < nn; ++k)
        for (int c = 0; c < 2; ++c) {
            int32_t ch = f.ref_nodes[k * 16 + 12 + c];
            if (ch >= 0) {
                if (ch % 4 != 0 || (size_t)(ch / 4) >= nn) return fail(TT_ERR_IO, "bvh cache: inner child out of range");
                f.dev_nodes[k * 16 + 12 + c] = ch / 4;
                f.dev_nodes[k * 16 + 14 + c] = 0;
            } else {
                offs.push_back(~ch);
            }
        }
    std::sort(offs.begin(), offs.end());
    offs.erase(std::unique(offs.begin(), offs.end()), offs.end());
    std::unordered_map<int32_t, std::pair<int32_t, int32_t>> leaf;  // vertex offset -> (first record, count)
    int32_t tri = 0;
    for (int32_t off : offs) {
        int32_t start = tri, cnt = 0;
        size_t v = (size_t)off;
        for (;;) {
            if (v >= nvert) return fail(TT_ERR_IO, "bvh cache: leaf runs past the vertex array");
            int32_t x;
            std::memcpy(&x, &f.ref_vertices[v * 4], 4);
            if (x == (int32_t)0x80000000) break;
            if (v + 2 >= nvert || v >= f.ref_tri_idx.size() || v >= f.ref_mat_idx.size())
                return fail(TT_ERR_IO, "bvh cache: truncated triangle");
            const float *p0 = &f.ref_vertices[v * 4], *p1 = p0 + 4, *p2 = p0 + 8;
            float id, mat;
            std::memcpy(&id, &f.ref_tri_idx[v], 4);
            std::memcpy(&mat, &f.ref_mat_idx[v], 4);
            float rec[12] = {p0[0], p0[1], p0[2], id, p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2], mat,
                             p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2], 0.0f};
            f.dev_tris.insert(f.dev_tris.end(), rec, rec + 12);
            ++tri;
            ++cnt;
            v += 3;
        }
        if (cnt == 0) {  // empty leaf: one degenerate record
            float rec[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            int32_t neg = -1;
            std::memcpy(&rec[3], &neg, 4);
            f.dev_tris.insert(f.dev_tris.end(), rec, rec + 12);
            ++tri;
        }
        int32_t one = 1;  // last-triangle flag
        std::memcpy(&f.dev_tris[(size_t)(tri - 1) * 12 + 11], &one, 4);
        leaf.emplace(off, std::make_pair(start, cnt));
    }
    for (size_t k = 0; k < nn; ++k)
        for (int c = 0; c < 2; ++c) {
            int32_t ch = f.ref_nodes[k * 16 + 12 + c];
            if (ch >= 0) continue;
            const auto &lc = leaf[~ch];
            f.dev_nodes[k * 16 + 12 + c] = ~lc.first;
            f.dev_nodes[k * 16 + 14 + c] = lc.second;
        }
    pair_align(f.dev_nodes);
    return TT_OK;
}

tt_status tt_bvh_load(const char *path, tt_bvh **out) {
    TT_GUARD_BEGIN
    if (!path || !out) return fail(TT_ERR_INVALID, "tt_bvh_load: bad arguments");
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(TT_ERR_IO, std::string("tt_bvh_load: cannot open ") + path);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(fp, std::fclose);
    uint32_t hdr[7];
    if (!rd(fp, hdr, sizeof(hdr))) return fail(TT_ERR_IO, "Error reading BVH cache file: short header");
    // sizes must be consistent with the layout before allocating
    if (hdr[0] % 4 != 0 || hdr[3] != hdr[4] || hdr[3] != hdr[5] || hdr[3] != hdr[6] || hdr[3] > (1u << 30) ||
        hdr[0] > (1u << 30))
        return fail(TT_ERR_IO, "Error reading BVH cache file: inconsistent header");
    auto b = std::make_unique<tt_bvh>();
    Flat &f = b->b.flat;
    f.tri_count = hdr[1];
    f.leaf_count = hdr[2];
    f.ref_nodes.resize((size_t)hdr[0] * 4);
    f.ref_vertices.resize((size_t)hdr[3] * 4);
    f.ref_normals.resize((size_t)hdr[4] * 4);
    f.ref_tri_idx.resize(hdr[5]);
    f.ref_mat_idx.resize(hdr[6]);
    if (!rd(fp, f.ref_nodes.data(), f.ref_nodes.size() * 4) || !rd(fp, f.ref_vertices.data(), f.ref_vertices.size() * 4) ||
        !rd(fp, f.ref_normals.data(), f.ref_normals.size() * 4) || !rd(fp, f.ref_tri_idx.data(), f.ref_tri_idx.size() * 4) ||
        !rd(fp, f.ref_mat_idx.data(), f.ref_mat_idx.size() * 4))
        return fail(TT_ERR_IO, "Error reading BVH cache file: truncated arrays");
    char magic[4];
    if (rd(fp, magic, 4) && std::memcmp(magic, "TTDL", 4) == 0) {
        uint32_t ver;
        int64_t dn, dt;
        if (!rd(fp, &ver, 4) || ver != 1 || !rd(fp, &dn, 8) || !rd(fp, &dt, 8) || dn < 0 || dt < 0 ||
            dn < (int64_t)hdr[0] / 4 || dn > 2 * ((int64_t)hdr[0] / 4) + 2 || dt > (int64_t)hdr[3])
            return fail(TT_ERR_IO, "Error reading BVH cache file: bad trailer");
        if (!rd(fp, &b->b.stats, sizeof(tt_bvh_stats))) return fail(TT_ERR_IO, "Error reading BVH cache file: bad trailer");
        f.dev_nodes.resize((size_t)dn * 16);
        f.dev_tris.resize((size_t)dt * 12);
        if (!rd(fp, f.dev_nodes.data(), f.dev_nodes.size() * 4) || !rd(fp, f.dev_tris.data(), f.dev_tris.size() * 4))
            return fail(TT_ERR_IO, "Error reading BVH cache file: truncated trailer");
    } else {
        tt_status st = compact_to_device(f);
        if (st != TT_OK) return st;
        b->b.stats = tt_bvh_stats{};
        b->b.stats.num_inner = (int32_t)(f.dev_nodes.size() / 16);
        b->b.stats.num_leaves = (int32_t)f.leaf_count;
        b->b.stats.num_refs = (int32_t)(f.dev_tris.size() / 12);
    }
    tt_status vs = check_device_layout(f);
    if (vs != TT_OK) return vs;
    *out = b.release();
    return TT_OK;
    TT_GUARD_END
}

}  // extern "C"
