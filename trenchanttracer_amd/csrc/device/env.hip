// Environment map upload and its importance-sampling distribution: per-texel
// weight = luminance * sin(theta) (equal-area correction of the equirectangular
// map). The GPU builds the conditional CDF per row and the marginal CDF over rows
// from FP64 block prefix sums (stored normalised as float, exported by
// tt_env_export_cdf); the host builds, from the same FP64 weights, the alias
// tables the renderer samples with (Vose's construction of Walker's method: a row
// from the row sums, then a column of that row, each one table load).
// Mapping (the reference's hdrSample, include/hdr/HDRKernel.cuh:16-27):
//   u = atan2(d.x, d.z) / 2pi wrapped to [0, 1), v = acos(d.y) / pi.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../host/tt_host.h"
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

}  // namespace

// Alias table of n weights (all >= 0) into tab[k] = {probability of keeping k,
// alias index as int bits}: Vose, "A linear algorithm for generating random
// numbers with a given distribution" (1991). Entry k is drawn with probability
// (prob[k] + sum of (1 - prob[m]) over the m aliased to k) / n = w[k] / sum(w).
// All-zero weights give the uniform table.
static void build_alias(const double *w, int n, float2 *tab, std::vector<double> &q, std::vector<int> &small,
                        std::vector<int> &large) {
    double sum = 0.0;
    for (int k = 0; k < n; ++k) sum += w[k];
    q.resize(n);
    small.clear();
    large.clear();
    for (int k = 0; k < n; ++k) {
        q[k] = sum > 0.0 ? w[k] * (double)n / sum : 1.0;
        (q[k] < 1.0 ? small : large).push_back(k);
    }
    auto put = [&](int k, double p, int a) {
        tab[k].x = (float)p;
        std::memcpy(&tab[k].y, &a, sizeof(int));
    };
    while (!small.empty() && !large.empty()) {
        const int l = small.back();
        small.pop_back();
        const int g = large.back();
        put(l, q[l], g);
        q[g] = (q[g] + q[l]) - 1.0;
        if (q[g] < 1.0) {
            large.pop_back();
            small.push_back(g);
        }
    }
    // what is left holds probability 1 up to rounding
    for (int k : large) put(k, 1.0, k);
    for (int k : small) put(k, 1.0, k);
}

static tt_status build_env_alias(tt_scene *s, const float *rgb, int32_t w, int32_t h) {
    std::vector<float2> cond((size_t)w * h), marg(h);
    std::vector<double> wt(w), rows(h), q;
    std::vector<int> small, large;
    for (int32_t row = 0; row < h; ++row) {
        const double sin_t = std::sin(3.14159265358979323846 * ((double)row + 0.5) / (double)h);
        double sum = 0.0;
        for (int32_t i = 0; i < w; ++i) {
            const float *p = rgb + 3 * ((size_t)row * w + i);
            const double lum = 0.2126 * (double)p[0] + 0.7152 * (double)p[1] + 0.0722 * (double)p[2];
            wt[i] = lum > 0.0 ? lum * sin_t : 0.0;
            sum += wt[i];
        }
        rows[row] = sum;
        build_alias(wt.data(), w, cond.data() + (size_t)row * w, q, small, large);
    }
    build_alias(rows.data(), h, marg.data(), q, small, large);
    tt_status st = dev_alloc(s->ctx, (void **)&s->env_alias_cond, cond.size() * sizeof(float2), "env alias rows");
    if (st == TT_OK) st = dev_alloc(s->ctx, (void **)&s->env_alias_marg, marg.size() * sizeof(float2), "env alias marginal");
    if (st != TT_OK) return st;
    hipError_t e = hipMemcpy(s->env_alias_cond, cond.data(), cond.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(s->env_alias_marg, marg.data(), marg.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("env alias upload: ") + hipGetErrorString(e));
    s->bytes += (cond.size() + marg.size()) * sizeof(float2);
    return TT_OK;
}

}  // namespace tt

extern "C" tt_status tt_debug_alias_table(const double *weights, int32_t n, float *out) {
    if (!weights || !out || n <= 0) return tt::fail(TT_ERR_INVALID, "tt_debug_alias_table: bad arguments");
    for (int32_t k = 0; k < n; ++k)
        if (!(weights[k] >= 0.0)) return tt::fail(TT_ERR_INVALID, "tt_debug_alias_table: weights must be >= 0");
    std::vector<float2> tab(n);
    std::vector<double> q;
    std::vector<int> small, large;
    tt::build_alias(weights, n, tab.data(), q, small, large);
    for (int32_t k = 0; k < n; ++k) {
        int a;
        std::memcpy(&a, &tab[k].y, sizeof(int));
        out[2 * k] = tab[k].x;
        out[2 * k + 1] = (float)a;
    }
    return TT_OK;
}

namespace tt {

tt_status build_env(tt_scene *s, const float *rgb, int32_t w, int32_t h, float scale) {
    tt_ctx *ctx = s->ctx;
    void *d_rgb = nullptr, *d_rows = nullptr;
    tt_status st = dev_alloc(ctx, &d_rgb, (size_t)w * h * 3 * sizeof(float), "env staging");
    if (st == TT_OK) st = dev_alloc(ctx, &d_rows, (size_t)h * sizeof(double), "env row sums");
    if (st == TT_OK) st = dev_alloc(ctx, (void **)&s->env_rgb, (size_t)w * h * sizeof(float4), "env texels");
    if (st == TT_OK) st = dev_alloc(ctx, (void **)&s->env_cond_cdf, (size_t)h * (w + 1) * sizeof(float), "env conditional cdf");
    if (st == TT_OK) st = dev_alloc(ctx, (void **)&s->env_marg_cdf, (size_t)(h + 1) * sizeof(float), "env marginal cdf");
    if (st == TT_OK) st = dev_alloc(ctx, (void **)&s->env_pdf_scale, sizeof(double), "env total");
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
    double total = 0.0;
    if (st == TT_OK && e == hipSuccess)
        e = hipMemcpyAsync(&total, s->env_pdf_scale, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (st == TT_OK && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (d_rgb) hipFree(d_rgb);
    if (d_rows) hipFree(d_rows);
    if (st != TT_OK) return st;
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("env build: ") + hipGetErrorString(e));
    st = build_env_alias(s, rgb, w, h);
    if (st != TT_OK) return st;
    s->env_w = w;
    s->env_h = h;
    s->env_scale = scale;
    s->env_total = total;
    s->bytes += (size_t)w * h * sizeof(float4) + (size_t)h * (w + 1) * 4 + (size_t)(h + 1) * 4 + 8;
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
