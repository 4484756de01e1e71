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
};

hipError_t guide_forward(tt_guide *g, int64_t n, const int32_t *n_dev, hipStream_t s);
hipError_t guide_train_rows(tt_guide *g, int64_t n, hipStream_t s);
tt_status guide_alloc_batch(tt_guide *g, int64_t rows);

}  // namespace tt

struct tt_guide {
    tt_ctx *ctx = nullptr;
    float *params = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
    float *scalars = nullptr;  // [0] loss, [1] sum of weights
    __bf16 *wf[4] = {nullptr, nullptr, nullptr, nullptr};  // forward copies Bt[out][in]
    __bf16 *wb[4] = {nullptr, nullptr, nullptr, nullptr};  // backward copies W[in][out_pad]
    float *partial = nullptr;
    int64_t partial_cap = 0;
    tt::GuideBatch batch;
    int64_t cap = 0;
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    int64_t step = 0;
    // online training from the renderer
    float *train_buf = nullptr;   // [cap][16] samples: inputs (9), dir (3), weight (1)
    int32_t *train_count = nullptr;
    int64_t train_cap = 0;
    float *rec = nullptr;         // per training path vertex records
    int64_t rec_cap = 0;
};
