// Host-side robustness driver for the library's host code (mesh generators, SBVH
// build, cache save / load, RGBE, OBJ and JSON scene loaders). It is linked from the
// host sources alone with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitized.py, which feeds it valid files and corrupted copies of
// them: every call must return a status (an error for a bad file) without a memory
// error, an out-of-bounds access or undefined behaviour.
//
//   host_fuzz gen <dir>                  build, save and reload a BVH, write an HDR
//   host_fuzz load <bvh|hdr|obj|config> <file>...
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tt.h"

#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "check failed: %s (%s)\n", #c, tt_last_error()); \
            return 1;                                                         \
        }                                                                     \
    } while (0)

// Copies out everything a loaded BVH exposes, so that every array it holds is read.
static int exercise_bvh(const tt_bvh *b) {
    tt_bvh_stats st;
    CHECK(tt_bvh_get_stats(b, &st) == TT_OK);
    tt_compact_sizes sz;
    CHECK(tt_bvh_export_compact(b, &sz, nullptr, nullptr, nullptr, nullptr, nullptr) == TT_OK);
    std::vector<int32_t> nodes((size_t)sz.nodes * 4 + 4), tri((size_t)sz.tri_indices + 1), mat((size_t)sz.mat_indices + 1);
    std::vector<float> verts((size_t)sz.vertices * 4 + 4), norms((size_t)sz.normals * 4 + 4);
    CHECK(tt_bvh_export_compact(b, &sz, nodes.data(), verts.data(), norms.data(), tri.data(), mat.data()) == TT_OK);
    int64_t nn = 0, nt = 0;
    CHECK(tt_bvh_export_device_layout(b, &nn, &nt, nullptr, nullptr) == TT_OK);
    std::vector<int32_t> dn((size_t)nn * 16 + 16);
    std::vector<float> dt((size_t)nt * 12 + 12);
    CHECK(tt_bvh_export_device_layout(b, &nn, &nt, dn.data(), dt.data()) == TT_OK);
    return 0;
}

static int gen(const std::string &dir) {
    tt_mesh *m = nullptr;
    CHECK(tt_mesh_gen_soup(3000, 7, -4.0f, 0.5f, &m) == TT_OK);
    tt_bvh *b = nullptr;
    CHECK(tt_bvh_build(m, 1e-5f, 2, &b) == TT_OK);
    CHECK(exercise_bvh(b) == 0);
    const std::string cache = dir + "/bvh.cache";
    CHECK(tt_bvh_save(b, cache.c_str()) == TT_OK);
    tt_bvh *b2 = nullptr;
    CHECK(tt_bvh_load(cache.c_str(), &b2) == TT_OK);
    tt_bvh_stats s1, s2;
    CHECK(tt_bvh_get_stats(b, &s1) == TT_OK && tt_bvh_get_stats(b2, &s2) == TT_OK);
    CHECK(s1.num_refs == s2.num_refs && s1.num_inner == s2.num_inner && s1.num_leaves == s2.num_leaves);
    CHECK(exercise_bvh(b2) == 0);
    std::vector<tt_ray> rays(32 * 24 * 2 + 1000);
    const float eye[3] = {0.5f, 0.5f, -1.0f}, target[3] = {0.5f, 0.5f, 0.0f};
    CHECK(tt_rays_primary(32, 24, 2, eye, target, 45.0f, 0.0f, 1e20f, 3, rays.data()) == TT_OK);
    CHECK(tt_rays_incoherent(m, 1000, 1e-4f, 1e20f, 3, rays.data() + 32 * 24 * 2) == TT_OK);
    tt_bvh_destroy(b2);
    tt_bvh_destroy(b);
    tt_mesh_destroy(m);
    // RGBE round trip of a small gradient (RGBE keeps 8 mantissa bits per channel)
    const int w = 17, h = 9;
    std::vector<float> img((size_t)w * h * 3);
    for (int i = 0; i < w * h * 3; ++i) img[i] = 0.01f + (float)(i % 97) * 0.37f;
    const std::string hdr = dir + "/sky.hdr";
    CHECK(tt_hdr_write(hdr.c_str(), w, h, img.data()) == TT_OK);
    int32_t rw = 0, rh = 0;
    CHECK(tt_hdr_read(hdr.c_str(), &rw, &rh, nullptr, 0) == TT_OK && rw == w && rh == h);
    std::vector<float> back((size_t)w * h * 3);
    CHECK(tt_hdr_read(hdr.c_str(), &rw, &rh, back.data(), (int64_t)back.size()) == TT_OK);
    // one shared exponent per pixel: each channel within 1% of the pixel's largest
    for (int px = 0; px < w * h; ++px) {
        const float mx = std::fmax(img[3 * px], std::fmax(img[3 * px + 1], img[3 * px + 2]));
        for (int c = 0; c < 3; ++c) CHECK(std::fabs(back[3 * px + c] - img[3 * px + c]) <= 0.01f * mx);
    }
    std::printf("gen ok\n");
    return 0;
}

static int load_one(const std::string &kind, const char *path, const char *base) {
    tt_status st = TT_OK;
    if (kind == "bvh") {
        tt_bvh *b = nullptr;
        st = tt_bvh_load(path, &b);
        if (st == TT_OK) {
            if (exercise_bvh(b)) return 1;
            tt_bvh_destroy(b);
        }
    } else if (kind == "hdr") {
        int32_t w = 0, h = 0;
        st = tt_hdr_read(path, &w, &h, nullptr, 0);
        if (st == TT_OK && (int64_t)w * h <= (1 << 22)) {
            std::vector<float> img((size_t)w * h * 3 + 3);
            st = tt_hdr_read(path, &w, &h, img.data(), (int64_t)w * h * 3);
        }
    } else if (kind == "obj") {
        tt_mesh *m = nullptr;
        int32_t nm = 0;
        st = tt_obj_load(path, &m, nullptr, 0, &nm);
        if (st == TT_OK) {
            tt_mesh_destroy(m);
            m = nullptr;
            std::vector<tt_material> mats((size_t)nm + 1);
            st = tt_obj_load(path, &m, mats.data(), nm, &nm);
            if (st == TT_OK) {
                int32_t nv = 0, nt = 0, nn = 0;
                CHECK(tt_mesh_info(m, &nv, &nt, &nn) == TT_OK);
                std::vector<float> v((size_t)nv * 3 + 3), n((size_t)nn * 3 + 3);
                std::vector<int32_t> t((size_t)nt * 3 + 3), tn((size_t)nt * 3 + 3), tm((size_t)nt + 1);
                CHECK(tt_mesh_get(m, v.data(), t.data(), n.data(), tn.data(), tm.data()) == TT_OK);
                tt_mesh_destroy(m);
            }
        }
    } else if (kind == "config") {
        tt_config *c = nullptr;
        st = tt_config_load(path, base, &c);
        if (st == TT_OK) {
            tt_config_info info;
            CHECK(tt_config_get(c, &info, nullptr, nullptr, nullptr) == TT_OK);
            std::vector<tt_material> mats((size_t)info.num_materials + 1);
            std::vector<tt_geometry> geoms((size_t)info.num_geometries + 1);
            tt_mesh *m = nullptr;
            st = tt_config_get(c, nullptr, mats.data(), geoms.data(), &m);
            if (m) tt_mesh_destroy(m);
            tt_config_destroy(c);
        }
    } else {
        std::fprintf(stderr, "unknown kind %s\n", kind.c_str());
        return 2;
    }
    std::printf("%s %s %d\n", kind.c_str(), path, (int)st);
    return 0;
}

int main(int argc, char **argv) {
    if (argc >= 3 && std::strcmp(argv[1], "gen") == 0) return gen(argv[2]);
    if (argc >= 4 && std::strcmp(argv[1], "load") == 0) {
        // files resolve their relative references against the first file's folder
        std::string base = argv[3];
        const size_t cut = base.find_last_of('/');
        base = cut == std::string::npos ? "." : base.substr(0, cut);
        for (int i = 3; i < argc; ++i)
            if (load_one(argv[2], argv[i], base.c_str())) return 1;
        return 0;
    }
    std::fprintf(stderr, "usage: host_fuzz gen <dir> | host_fuzz load <bvh|hdr|obj|config> <file>...\n");
    return 2;
}
