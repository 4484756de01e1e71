// Guiding network: sizes and device objects.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/tt.h"

namespace tt {

constexpr int kGuideIn = 9;          // position (3), normal (3), outgoing direction (3), each in [0, 1]
constexpr int kGuideFreq = 12;       // sin / cos (2^f pi x), f = 0..11
constexpr int kGuideFeat = kGuideIn * kGuideFreq * 2;  // 216
constexpr int kGuideK1 = 224;        // features padded to a multiple of 32
constexpr int kGuideH = 64;          // hidden width (3 hidden layers)
constexpr int kGuideLobes = 8;
constexpr int kGuideNOut = kGuideLobes * 5;  // 40: mean (3), log kappa (1), logit (1) per lobe
// output 40 (in the padding of the last layer): the logit z of the learned selection
// probability of the mixture, alpha = kAlphaLo + (kAlphaHi - kAlphaLo) sigmoid(z)
constexpr int kGuideAlphaOut = kGuideNOut;
constexpr float kAlphaLo = 0.02f, kAlphaHi = 0.95f;
constexpr int kGuideNOutPad16 = 48;  // output columns computed (multiple of 16); raw row stride
constexpr int kGuideNOutPad = 64;    // output-gradient width (multiple of 32, the backward K)
constexpr float kGuideLogKappaMin = -5.0f, kGuideLogKappaMax = 9.0f;

struct GuideLayer {
    int in, out;
};
extern const GuideLayer kGuideLayers[4];
int64_t guide_param_count();

struct GuideBatch {
    float *in = nullptr;      // [R][9]
    __bf16 *x = nullptr;      // [R][224] features
    __bf16 *h[3] = {nullptr, nullptr, nullptr};   // [R][64]
    float *raw = nullptr;     // [R][48] outputs
    __bf16 *draw = nullptr;   // [R][64] output gradients
    __bf16 *dh[3] = {nullptr, nullptr, nullptr};  // [R][64]
    float *dirs = nullptr;    // [R][3] training directions
    float *wts = nullptr;     // [R] training weights (radiance / pdf)
    float *pb = nullptr;      // [R] BSDF density of each direction (learned alpha), or unset
    float *lpart = nullptr;   // per loss-gradient block: loss sums, then weight sums (fixed-order adds)
};

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float learned_alpha(float z) { return kAlphaLo + (kAlphaHi - kAlphaLo) * sigmoidf_(z); }

// ------------------------------------------------------------------ vMF mixture
// Render-side form of the mixture: per lobe the unit mean, kappa, the weight (for
// picking a lobe) and lc = log(weight) + log of the vMF normaliser, so that the
// density at a unit direction is sum_k exp(lc_k + kappa_k (mu_k . w - 1)) — one exp
// per lobe per evaluation instead of a log, a log1p and two exps. Decoded once per
// vertex; the density is evaluated for the next-event direction and the sampled one.
struct Mixture {
    float mu[kGuideLobes][3];
    float kappa[kGuideLobes];
    float w[kGuideLobes];
    float lc[kGuideLobes];
};

// The render-side decode and density use the hardware approximations of exp, log,
// sqrt and reciprocal (about 1e-6 relative error, each a single instruction where
// the IEEE-exact forms take sequences of 10-30): the one-sample MIS weight only needs
// the density the sampler draws from to float precision, and both use these same
// functions. log(1 - e^{-2 kappa}) is taken directly: at the smallest kappa (e^-5)
// the difference 1 - e^{-2 kappa} = 0.0133 keeps a relative error near 1e-5.
__device__ __forceinline__ void decode_mixture(const float *raw, Mixture &M) {
    float mx = -1e30f;
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) mx = fmaxf(mx, raw[kGuideLobes * 4 + k]);
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        const float m0 = raw[3 * k], m1 = raw[3 * k + 1], m2 = raw[3 * k + 2];
        const float inv = __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(m0 * m0 + m1 * m1 + m2 * m2) + 1e-6f);
        M.mu[k][0] = m0 * inv;
        M.mu[k][1] = m1 * inv;
        M.mu[k][2] = m2 * inv;
        const float lk = fminf(fmaxf(raw[kGuideLobes * 3 + k], kGuideLogKappaMin), kGuideLogKappaMax);
        M.kappa[k] = __expf(lk);
        // log of kappa / (2 pi (1 - e^-2k)), with log kappa taken from the output itself
        M.lc[k] = lk - 1.83787706641f - __logf(1.0f - __expf(-2.0f * M.kappa[k]));
        const float d = raw[kGuideLobes * 4 + k] - mx;
        M.w[k] = __expf(d);
        M.lc[k] += d;
        sum += M.w[k];
    }
    const float lsum = __logf(sum), isum = __builtin_amdgcn_rcpf(sum);
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        M.w[k] *= isum;
        M.lc[k] -= lsum;
    }
}

__device__ __forceinline__ float mixture_density(const Mixture &M, float x, float y, float z) {
    float q = 0.0f;
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        const float c = M.mu[k][0] * x + M.mu[k][1] * y + M.mu[k][2] * z;
        q += __expf(M.lc[k] + M.kappa[k] * (c - 1.0f));
    }
    return q;
}

// Samples the mixture: lobe by cumulative weight (u0), then the vMF lobe by
// Jakob's stable inversion w = 1 + log(u1 + (1 - u1) e^{-2k}) / k, phi = 2 pi u2.
template <class Mix>
__device__ __forceinline__ void mixture_sample(const Mix &L, float u0, float u1, float u2, float *out) {
    // the lobe is picked with constant indices only (selects, no break), so the
    // mixture stays in registers
    float kap = L.kappa[kGuideLobes - 1], mx = L.mu[kGuideLobes - 1][0], my = L.mu[kGuideLobes - 1][1],
          mz = L.mu[kGuideLobes - 1][2];
    float acc = 0.0f;
    bool found = false;
#pragma unroll
    for (int j = 0; j < kGuideLobes; ++j) {
        acc += L.w[j];
        const bool take = !found && u0 <= acc;
        kap = take ? L.kappa[j] : kap;
        mx = take ? L.mu[j][0] : mx;
        my = take ? L.mu[j][1] : my;
        mz = take ? L.mu[j][2] : mz;
        found = found || take;
    }
    const float w = fminf(1.0f, fmaxf(-1.0f, 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * kap)) / kap));
    const float sr = sqrtf(fmaxf(0.0f, 1.0f - w * w));
    float sp, cp;
    __sincosf(6.28318531f * u2, &sp, &cp);
    // frame around mu (Duff et al. 2017)
    const float s = mz >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (s + mz);
    const float b = mx * my * a;
    const float tx = 1.0f + s * mx * mx * a, ty = s * b, tz = -s * mx;
    const float bx = b, by = s + my * my * a, bz = -my;
    const float lx = sr * cp, ly = sr * sp;
    float dx = tx * lx + bx * ly + mx * w, dy = ty * lx + by * ly + my * w, dz = tz * lx + bz * ly + mz * w;
    const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    out[0] = dx * inv;
    out[1] = dy * inv;
    out[2] = dz * inv;
}

hipError_t guide_forward(tt_guide *g, int64_t n, const int32_t *n_dev, hipStream_t s);
// slot_of (may be null): row r of the inference reads inputs and writes outputs at
// row slot_of[r] of in and raw (a compacted list of the rows needed); rows up to
// min(*n_dev, n). store_acts: also write the features and hidden activations into
// g->batch (training)
hipError_t guide_infer(tt_guide *g, const float *in, float *raw, int64_t n, const int32_t *n_dev, hipStream_t s,
                       const int32_t *slot_of = nullptr, bool store_acts = false);
hipError_t guide_train_rows(tt_guide *g, int64_t n, hipStream_t s);
tt_status guide_alloc_batch(tt_guide *g, int64_t rows);
tt_status guide_alloc_infer(tt_guide *g, int64_t rows);

}  // namespace tt

struct tt_guide {
    tt_ctx *ctx = nullptr;
    float *params = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
    float *scalars = nullptr;  // [0] loss, [1] sum of the training weights
    __bf16 *wf[4] = {nullptr, nullptr, nullptr, nullptr};  // forward copies Bt[out][in]
    __bf16 *wb[4] = {nullptr, nullptr, nullptr, nullptr};  // backward copies W[in][out_pad]
    float *partial = nullptr;
    int64_t partial_cap = 0;
    tt::GuideBatch batch;
    int64_t cap = 0;
    // renderer inference rows (one per queued path): inputs [icap][9], mixtures [icap][48]
    float *iin = nullptr, *iraw = nullptr;
    uint64_t *imask = nullptr;    // one bit per queue slot: its vertex is guided (its row is inferred)
    int32_t *iblock_n = nullptr;  // guided slots of each block of the guide-row kernels; [iblocks]: the row total
    int32_t *islot = nullptr;     // the guided slots in slot order (the inference's rows)
    int64_t icap = 0, iblocks = 0;
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    int64_t step = 0;
    // online training from the renderer
    float *train_buf = nullptr;   // [cap][16] samples: inputs (9), dir (3), weight (1), BSDF density (1)
    int32_t *train_off = nullptr;   // per block of 256 records: row offset; then the row total
    int64_t train_off_cap = 0;
    int64_t train_cap = 0;
    float *rec = nullptr;         // per training path vertex records
    int64_t rec_cap = 0;
    bool train_alpha = false;     // the staged batch carries BSDF densities: train the alpha output
};
