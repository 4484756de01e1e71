// Device objects behind the C ABI handles.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../../include/tt.h"

namespace tt {
struct EnvCell;
struct MargCell;
}

struct tt_ctx {
    int device = 0;
    int num_cus = 256;
    char arch[64] = {0};
    hipStream_t stream = nullptr;
    void *scratch = nullptr;  // staging for the host-buffer wrappers
    size_t scratch_bytes = 0;
    void *reorder_ws = nullptr;  // ray-reordering keys, permutation and sort storage
    size_t reorder_ws_bytes = 0;
    hipEvent_t reorder_done = nullptr;  // recorded after the last launch that used reorder_ws
    bool reorder_pending = false;       // reorder_done has been recorded at least once
    // path-mode renders: each bounce's next-event work (shadow trace, occlusion, apply)
    // runs on aux, beside the next bounce's extension trace on stream (default priority)
    hipStream_t aux = nullptr;
    hipEvent_t shade_done = nullptr;    // recorded on stream after a bounce's shade kernel
    hipEvent_t shadow_done[2] = {nullptr, nullptr};   // on aux after the next-event work of even / odd bounces
};

struct tt_scene {
    tt_ctx *ctx = nullptr;
    // SBVH device layout
    void *nodes = nullptr;       // float4 x 4 per node
    void *tris = nullptr;        // float4 x 3 per triangle record
    int64_t num_nodes = 0;
    int64_t num_tris = 0;
    // shading data by original triangle id
    void *tri_normals = nullptr;  // float4 x 3 per triangle
    int32_t *tri_mat = nullptr;
    int32_t num_orig_tris = 0;
    float *tri_geo = nullptr;     // float4 per triangle: unit geometric normal (e1 x e2), area
    uint32_t *shade_tri = nullptr; // 32 B per triangle for path-mode shading: geometric normal (3 floats),
                                   // material id, the 3 vertex normals octahedral-encoded (2 x 16 bits), pad
    tt_material *mats = nullptr;
    int32_t num_mats = 0;
    // emissive triangles (area-proportional light sampling)
    int32_t *emit_tri = nullptr;  // original triangle ids
    float *emit_cdf = nullptr;    // num_emit + 1
    float *emit_verts = nullptr;  // float4 x 3 per emitter
    int32_t num_emit = 0;
    float emit_area = 0.0f;
    // environment map (equirectangular) and its sampling distribution
    float *env_rgb = nullptr;        // float4 per texel (rgb, luminance weight)
    float *env_marg_cdf = nullptr;   // H + 1 floats
    float *env_cond_cdf = nullptr;   // H * (W + 1) floats
    tt::MargCell *env_marg_cells = nullptr;  // H cells of the marginal CDF's guide table
    tt::EnvCell *env_cells = nullptr;    // H * W cells of the conditional CDFs' guide tables, with their texels
    double *env_pdf_scale = nullptr; // [0] = total weight (FP64)
    int32_t env_w = 0, env_h = 0;
    float env_scale = 1.0f;
    double env_total = 0.0;
    float bmin[3] = {0.0f, 0.0f, 0.0f}, bmax[3] = {1.0f, 1.0f, 1.0f};  // mesh bounds (guide positions)
    tt_geometry *geoms = nullptr;  // analytic primitives (device)
    int32_t num_geoms = 0;
    size_t bytes = 0;
    struct RenderState *rs = nullptr;  // wavefront buffers, sized on first render
};

struct tt_timer {
    tt_ctx *ctx = nullptr;
    std::vector<hipEvent_t> ev;
};

namespace tt {
tt_status dev_alloc(tt_ctx *ctx, void **p, size_t bytes, const char *what);
tt_status build_env(tt_scene *s, const float *rgb, int32_t w, int32_t h, float scale);
void free_render_state(tt_scene *s);
}  // namespace tt
