// Environment map upload and its importance-sampling distribution, built on the
// GPU: per-texel weight = luminance * sin(theta) (equal-area correction of the
// equirectangular map), conditional CDF per row and marginal CDF over rows from
// FP64 block prefix sums, stored normalised as float, and a guide table per CDF
// (one cell per entry) that confines each inversion to a short index range.
// Mapping (the reference's hdrSample, include/hdr/HDRKernel.cuh:16-27):
//   u = atan2(d.x, d.z) / 2pi wrapped to [0, 1), v = acos(d.y) / pi.
#include <algorithm>
#include <string>
#include <vector>

#include "../host/tt_host.h"
#include "dev_common.h"
#include "scene_dev.h"

namespace tt {
namespace {

constexpr int kScanThreads = 256;

// Exclusive block scan of one double per thread (Hillis–Steele in LDS).
__device__ double block_scan_excl(double x, double *sh, double &total) {
    const int t = threadIdx.x;
    sh[t] = x;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        double y = t >= off ? sh[t - off] : 0.0;
        __syncthreads();
        sh[t] += y;
        __syncthreads();
    }
    total = sh[kScanThreads - 1];
    double incl = sh[t];
    __syncthreads();
    return incl - x;
}

__global__ __launch_bounds__(kScanThreads) void k_env_rows(const float *__restrict__ rgb, float4 *__restrict__ texel,
                                                           float *__restrict__ cond, double *__restrict__ row_sum,
                                                           int w, int h) {
    __shared__ double sh[kScanThreads];
    const int row = blockIdx.x;
    const int per = (w + kScanThreads - 1) / kScanThreads;
    const int lo = min(w, (int)threadIdx.x * per), hi = min(w, lo + per);
    const double sin_t = sin(3.14159265358979323846 * ((double)row + 0.5) / (double)h);
    double local = 0.0;
    for (int i = lo; i < hi; ++i) {
        const float *p = rgb + 3 * ((size_t)row * w + i);
        double lum = 0.2126 * (double)p[0] + 0.7152 * (double)p[1] + 0.0722 * (double)p[2];
        double wt = lum > 0.0 ? lum * sin_t : 0.0;
        texel[(size_t)row * w + i] = make_float4(p[0], p[1], p[2], (float)wt);
        local += wt;
    }
    double total;
    double run = block_scan_excl(local, sh, total);
    float *c = cond + (size_t)row * (w + 1);
    if (threadIdx.x == 0) c[0] = 0.0f;
    for (int i = lo; i < hi; ++i) {
        const float *p = rgb + 3 * ((size_t)row * w + i);
        double lum = 0.2126 * (double)p[0] + 0.7152 * (double)p[1] + 0.0722 * (double)p[2];
        run += lum > 0.0 ? lum * sin_t : 0.0;
        c[i + 1] = total > 0.0 ? (float)(run / total) : (float)(i + 1) / (float)w;
    }
    if (threadIdx.x == 0) row_sum[row] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_env_marginal(const double *__restrict__ row_sum,
                                                               float *__restrict__ marg, double *__restrict__ total_out,
                                                               int h) {
    __shared__ double sh[kScanThreads];
    const int per = (h + kScanThreads - 1) / kScanThreads;
    const int lo = min(h, (int)threadIdx.x * per), hi = min(h, lo + per);
    double local = 0.0;
    for (int j = lo; j < hi; ++j) local += row_sum[j];
    double total;
    double run = block_scan_excl(local, sh, total);
    if (threadIdx.x == 0) marg[0] = 0.0f;
    for (int j = lo; j < hi; ++j) {
        run += row_sum[j];
        marg[j + 1] = total > 0.0 ? (float)(run / total) : (float)(j + 1) / (float)h;
    }
    if (threadIdx.x == 0) total_out[0] = total;
}

// Smallest float u in (0, 1] whose cell (guide_cell's arithmetic, unclamped) is
// >= c, for c >= 1: bisection over the float bit patterns, which order
// nonnegative floats.
__device__ float cell_start(int c, int m) {
    uint32_t lo = 0u, hi = 0x3f800000u;   // cell(lo) < c <= cell(hi)
    while (hi - lo > 1u) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if ((int)(__uint_as_float(mid) * (float)m) >= c) hi = mid;
        else lo = mid;
    }
    return __uint_as_float(hi);
}

// Guide tables of `rows` CDFs of n entries each (row r at cdf + r * cdf_stride):
// cell c of a row gets the full search's index at the first and the last float u
// that guide_cell maps to c (the search is monotone in u, so every u of the cell
// lands between them).
__global__ __launch_bounds__(256) void k_env_guide(const float *__restrict__ cdf, int64_t cdf_stride, int n, int rows,
                                                   int2 *__restrict__ guide) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const float u_lo = c == 0 ? 0.0f : cell_start(c, n);
    const float u_hi = c == n - 1 ? 0x1.fffffep-1f : __uint_as_float(__float_as_uint(cell_start(c + 1, n)) - 1u);
    for (int r = blockIdx.y; r < rows; r += gridDim.y) {
        const float *cd = cdf + (int64_t)r * cdf_stride;
        const int lo = cdf_search_range(cd, 0, n - 1, u_lo);
        const int hi = u_hi < u_lo ? lo : cdf_search_range(cd, 0, n - 1, u_hi);
        guide[(int64_t)r * n + c] = make_int2(lo, hi);
    }
}

// Cell records of the rows' conditional guide tables: each cell's search range, the CDF
// entries and texels a sample drawn in the cell can need (EnvCell).
__global__ __launch_bounds__(256) void k_env_cells(const int2 *__restrict__ guide, const float *__restrict__ cond,
                                                   const float4 *__restrict__ texel, int w, int h,
                                                   EnvCell *__restrict__ cells) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w) return;
    for (int r = blockIdx.y; r < h; r += gridDim.y) {
        const int2 g = guide[(int64_t)r * w + c];
        const float *cd = cond + (int64_t)r * (w + 1);
        const float4 *tr = texel + (int64_t)r * w;
        EnvCell e;
        e.lo = g.x;
        e.hi = g.y;
        e.c1 = g.x + 1 <= w ? cd[g.x + 1] : 0.0f;
        e.c2 = g.x + 2 <= w ? cd[g.x + 2] : 0.0f;
        for (int k = 0; k < 3; ++k) e.t[k] = g.x + k < w ? tr[g.x + k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        cells[(int64_t)r * w + c] = e;
    }
}

// Records of the marginal CDF's guide cells (MargCell).
__global__ __launch_bounds__(256) void k_env_marg_cells(const int2 *__restrict__ guide, const float *__restrict__ marg,
                                                        int h, MargCell *__restrict__ cells) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= h) return;
    const int2 g = guide[c];
    MargCell e;
    e.lo = g.x;
    e.hi = g.y;
    e.c1 = g.x + 1 <= h ? marg[g.x + 1] : 0.0f;
    e.c2 = g.x + 2 <= h ? marg[g.x + 2] : 0.0f;
    cells[c] = e;
}

}  // namespace

tt_status build_env(tt_scene *s, const float *rgb, int32_t w, int32_t h, float scale) {
    tt_ctx *ctx = s->ctx;
    void *d_rgb = nullptr, *d_rows = nullptr, *d_guide = nullptr;
    tt_status st = dev_alloc(ctx, &d_rgb, (size_t)w * h * 3 * sizeof(float), "env staging");
    if (st == TT_OK) st = dev_alloc(ctx, &d_rows, (size_t)h * sizeof(double), "env row sums");
    if (st == TT_OK) st = dev_alloc(ctx, (void **)&s->env_rgb, (size_t)w * h * sizeof(float4), "env texels");
    if (st == TT_OK) st = dev_alloc(ctx, (void **)&s->env_cond_cdf, (size_t)h * (w + 1) * sizeof(float), "env conditional cdf");
    if (st == TT_OK) st = dev_alloc(ctx, (void **)&s->env_marg_cdf, (size_t)(h + 1) * sizeof(float), "env marginal cdf");
    if (st == TT_OK) st = dev_alloc(ctx, (void **)&s->env_pdf_scale, sizeof(double), "env total");
    if (st == TT_OK) st = dev_alloc(ctx, &d_guide, (size_t)h * w * sizeof(int2), "env conditional guide");
    if (st == TT_OK) st = dev_alloc(ctx, (void **)&s->env_cells, (size_t)h * w * sizeof(EnvCell), "env cell records");
    if (st == TT_OK) st = dev_alloc(ctx, (void **)&s->env_marg_cells, (size_t)h * sizeof(MargCell), "env marginal cells");
    hipError_t e = hipSuccess;
    if (st == TT_OK) e = hipMemcpy(d_rgb, rgb, (size_t)w * h * 3 * sizeof(float), hipMemcpyHostToDevice);
    if (st == TT_OK && e == hipSuccess) {
        hipLaunchKernelGGL(k_env_rows, dim3(h), dim3(kScanThreads), 0, ctx->stream, (const float *)d_rgb,
                           (float4 *)s->env_rgb, s->env_cond_cdf, (double *)d_rows, w, h);
        e = hipGetLastError();
    }
    if (st == TT_OK && e == hipSuccess) {
        hipLaunchKernelGGL(k_env_marginal, dim3(1), dim3(kScanThreads), 0, ctx->stream, (const double *)d_rows,
                           s->env_marg_cdf, s->env_pdf_scale, h);
        e = hipGetLastError();
    }
    if (st == TT_OK && e == hipSuccess) {
        hipLaunchKernelGGL(k_env_guide, dim3((unsigned)((w + 255) / 256), (unsigned)std::min(h, 65535)), dim3(256), 0,
                           ctx->stream, (const float *)s->env_cond_cdf, (int64_t)(w + 1), w, h, (int2 *)d_guide);
        e = hipGetLastError();
    }
    if (st == TT_OK && e == hipSuccess) {
        hipLaunchKernelGGL(k_env_cells, dim3((unsigned)((w + 255) / 256), (unsigned)std::min(h, 65535)), dim3(256), 0,
                           ctx->stream, (const int2 *)d_guide, (const float *)s->env_cond_cdf,
                           (const float4 *)s->env_rgb, w, h, s->env_cells);
        e = hipGetLastError();
    }
    if (st == TT_OK && e == hipSuccess) {
        hipLaunchKernelGGL(k_env_guide, dim3((unsigned)((h + 255) / 256), 1), dim3(256), 0, ctx->stream,
                           (const float *)s->env_marg_cdf, (int64_t)(h + 1), h, 1, (int2 *)d_guide);
        e = hipGetLastError();
    }
    if (st == TT_OK && e == hipSuccess) {
        hipLaunchKernelGGL(k_env_marg_cells, dim3((unsigned)((h + 255) / 256)), dim3(256), 0, ctx->stream,
                           (const int2 *)d_guide, (const float *)s->env_marg_cdf, h, s->env_marg_cells);
        e = hipGetLastError();
    }
    double total = 0.0;
    if (st == TT_OK && e == hipSuccess)
        e = hipMemcpyAsync(&total, s->env_pdf_scale, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (st == TT_OK && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (d_rgb) hipFree(d_rgb);
    if (d_rows) hipFree(d_rows);
    if (d_guide) hipFree(d_guide);
    if (st != TT_OK) return st;
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("env build: ") + hipGetErrorString(e));
    s->env_w = w;
    s->env_h = h;
    s->env_scale = scale;
    s->env_total = total;
    s->bytes += (size_t)w * h * sizeof(float4) + (size_t)h * (w + 1) * 4 + (size_t)(h + 1) * 4 + 8 +
                (size_t)h * sizeof(MargCell) + (size_t)h * w * sizeof(EnvCell);
    return TT_OK;
}

}  // namespace tt

extern "C" tt_status tt_env_export_cdf(const tt_scene *s, float *marg, float *cond, double *total) {
    if (!s || !s->env_rgb) return tt::fail(TT_ERR_INVALID, "tt_env_export_cdf: scene has no env map");
    hipError_t e = hipSuccess;
    if (marg) e = hipMemcpy(marg, s->env_marg_cdf, (size_t)(s->env_h + 1) * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && cond)
        e = hipMemcpy(cond, s->env_cond_cdf, (size_t)s->env_h * (s->env_w + 1) * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return tt::fail(TT_ERR_DEVICE, std::string("tt_env_export_cdf: ") + hipGetErrorString(e));
    if (total) *total = s->env_total;
    return TT_OK;
}
