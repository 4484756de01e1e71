// Learned path-guiding proposal: a small MLP trained online on the tracer's own
// radiance samples (no reference counterpart).
//
// Network: 9 inputs (position normalised to the scene box, shading normal and
// outgoing direction mapped to [0, 1]) -> frequency encoding sin / cos(2^f pi x),
// f = 0..11 (216 features, padded to 224) -> 3 hidden layers of 64 (ReLU) -> 40
// outputs (padded to 48): an 8-lobe von Mises-Fisher mixture, per lobe a raw mean
// direction (3), log concentration (1) and logit weight (1).
// BF16 weights and activations on MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation; FP32 master weights and Adam moments.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 on wave64 (lane l):
//   A: A[l % 16][8 (l / 16) + j], B: B[8 (l / 16) + j][l % 16], j = 0..7
//   D: D[4 (l / 16) + j][l % 16], j = 0..3
// Matrices are kept so that each lane's 8 operand values are 16 contiguous bytes:
// activations row-major [rows][K], weights as Bt[out][in] for the forward pass
// and as W[in][out] for the backward pass (both copies refreshed by Adam).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../common/philox.h"
#include "../host/tt_host.h"
#include "guide_dev.h"
#include "scene_dev.h"
#include "ttmath.h"

namespace tt {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------ encoding
// Frequency features of one input x: sin / cos(2^f pi x), f = 0..11. The base
// pair comes from the bit-reproducible tm_sincos(pi x); each next frequency from
// the double-angle recurrences s' = (2 s) c, c' = (c - s)(c + s) in float (no
// contraction), so the features are a fixed function of x that a host
// restatement reproduces bit for bit. Each doubling at most doubles the error
// (both maps have gradient norm <= 2), so after 11 doublings the absolute error
// stays near 1e-4, well under the BF16 rounding that follows.
struct FreqState {
    float s, c;
};
__device__ __forceinline__ FreqState freq_base(float x) {
    FreqState st;
    st.s = tm_sincos(x * 3.14159265f, &st.c);
    return st;
}
__device__ __forceinline__ void freq_next(FreqState &st) {
    const float s2 = (2.0f * st.s) * st.c;
    st.c = (st.c - st.s) * (st.c + st.s);
    st.s = s2;
}

// x[r][k], k = 24 c + 2 f + {0: sin, 1: cos}, zero for k >= 216 and rows >= n:
// one thread per (row, input); the 9th "input" writes the 8 padding features.
__global__ __launch_bounds__(256) void k_encode(const float *__restrict__ in, __bf16 *__restrict__ x, int64_t n,
                                                const int32_t *n_dev) {
    int64_t rows = n_dev ? min((int64_t)*n_dev, n) : n;
    const int64_t rows_pad = (rows + 15) & ~15LL;
    constexpr int kSlots = kGuideIn + 1;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < rows_pad * kSlots;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / kSlots;
        const int c = (int)(t % kSlots);
        __bf16 *o = x + r * kGuideK1 + 2 * kGuideFreq * c;
        if (c == kGuideIn) {
            for (int k = 0; k < kGuideK1 - kGuideFeat; ++k) o[k] = (__bf16)0.0f;
            continue;
        }
        FreqState st = freq_base(r < rows ? in[r * kGuideIn + c] : 0.0f);
        for (int f = 0; f < kGuideFreq; ++f) {
            o[2 * f] = (__bf16)(r < rows ? st.s : 0.0f);
            o[2 * f + 1] = (__bf16)(r < rows ? st.c : 0.0f);
            freq_next(st);
        }
    }
}

// ------------------------------------------------------------------ GEMM
// out[r][c] = act(sum_k A[r][k] Bt[c][k] + bias[c]), rows r < n (device count if
// n_dev), K % 32 == 0, NC % 16 == 0 (<= 64). One wave per 16-row tile.
// mask (optional, bf16 [rows][NC]): out *= (mask > 0) (ReLU derivative).
template <int K, int NC, bool RELU, bool OUT_BF16>
__global__ __launch_bounds__(256) void k_gemm(const __bf16 *__restrict__ A, const __bf16 *__restrict__ Bt,
                                              const float *__restrict__ bias, void *__restrict__ out, int ldo,
                                              const __bf16 *__restrict__ mask, int64_t n, const int32_t *n_dev) {
    constexpr int T = NC / 16;
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    int64_t rows = n_dev ? min((int64_t)*n_dev, n) : n;
    const int64_t tiles = (rows + 15) >> 4;
    const int lr = lane & 15, lq = lane >> 4;
    for (int64_t tile = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < tiles; tile += waves) {
        f32x4 acc[T];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        const __bf16 *arow = A + (tile * 16 + lr) * K + 8 * lq;
#pragma unroll 2
        for (int s = 0; s < K / 32; ++s) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(arow + 32 * s);
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(Bt + (int64_t)(16 * t + lr) * K + 32 * s + 8 * lq);
                acc[t] = mfma(a, b, acc[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int col = 16 * t + lr;
            const float bcol = bias ? bias[col] : 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t row = tile * 16 + 4 * lq + j;
                float v = acc[t][j] + bcol;
                if (RELU) v = fmaxf(v, 0.0f);
                if (mask && !((float)mask[row * NC + col] > 0.0f)) v = 0.0f;
                if (OUT_BF16) reinterpret_cast<__bf16 *>(out)[row * ldo + col] = (__bf16)v;
                else reinterpret_cast<float *>(out)[row * ldo + col] = v;
            }
        }
    }
}

// ------------------------------------------------------------------ fused inference
// raw[r] = MLP(encode(in[r])) for r < rows: encoding, 3 hidden layers and the
// output layer in one kernel. One block of kInferWaves waves per CU stages all
// BF16 weights (Bt[out][in], rows padded by 8 elements against LDS bank conflicts)
// and FP32 biases in LDS once; its waves then take 16-row tiles. A tile's features
// are computed once per (row, input) frequency chain (144 chains over 64 lanes)
// into the wave's LDS tile, read back in the MFMA A layout, and the same tile then
// carries each hidden layer's output (D layout -> A layout); the last layer writes
// FP32. Arithmetic, rounding points and accumulation order are those of k_encode +
// k_gemm, so the outputs are identical to the unfused forward pass.
constexpr int kInferWaves = 14;
constexpr int kLd1 = kGuideK1 + 8, kLdH = kGuideH + 8;
constexpr int kInferTile = 16 * kLd1;   // bf16 elements per wave tile (features, then activations)
constexpr int kInferLds = 2 * (kGuideH * kLd1 + 2 * kGuideH * kLdH + kGuideNOutPad16 * kLdH + kInferWaves * kInferTile) +
                          4 * (3 * kGuideH + kGuideNOutPad16);
static_assert(kInferLds <= 160 * 1024, "inference block exceeds the CU's LDS");
static_assert(16 * kLdH <= kInferTile, "activation tile must fit in the feature tile");
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

struct InferArgs {
    const float *in;
    const __bf16 *w[4];
    const float *bias[4];
    float *raw;
    int64_t n;
    const int32_t *n_dev;
    // if set: row r reads its inputs from and writes its outputs to row slot_of[r] of in
    // and raw (a compacted list of the rows needed); rows past the count write nothing
    const int32_t *slot_of;
    // training forward: if set, the three hidden layers' activations [rows][64] of every
    // 16-row tile are written too (the backward pass reads them), for all rows of the tile
    __bf16 *h_out[3];
};

// The wave's 16-row LDS tile (row stride LD) to global rows [tile * 16, tile * 16 + 16)
// of a row-major [rows][K] array, 16 bytes per lane.
template <int K, int LD>
__device__ __forceinline__ void tile_to_global(const __bf16 *tile_lds, __bf16 *dst, int64_t tile, int lane) {
    constexpr int VEC = 16 * K / 8;
    for (int v = lane; v < VEC; v += 64) {
        const int rr = v / (K / 8), c = v % (K / 8);
        *reinterpret_cast<bf16x8 *>(dst + (tile * 16 + rr) * K + 8 * c) =
            *reinterpret_cast<const bf16x8 *>(tile_lds + rr * LD + 8 * c);
    }
}

template <int K, int LD>
__device__ __forceinline__ void stage_weights(__bf16 *dst, const __bf16 *src, int rows) {
    // rows x K bf16 from global (row-major, K contiguous) into LDS rows of stride LD
    const int vec = rows * K / 8;
    for (int t = threadIdx.x; t < vec; t += blockDim.x) {
        const int r = t / (K / 8), c = t % (K / 8);
        *reinterpret_cast<bf16x8 *>(dst + r * LD + 8 * c) = *reinterpret_cast<const bf16x8 *>(src + (int64_t)r * K + 8 * c);
    }
}

// (A W^T)^T over K for a 16-row tile A (row-major in LDS, row stride LD) and W[out][K]
// (stride LD): W's rows are the MFMA's A operand and the tile's rows its B operand, so
// each output element is the dot product k_gemm computes, and the lane holds columns
// 16 t + 4 lq + j of tile row lr.
template <int NT, int K, int LD>
__device__ __forceinline__ void tile_mfma_t(const __bf16 *act, const __bf16 *w, int lr, int lq, f32x4 *acc) {
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < K / 32; ++s) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(act + lr * LD + 32 * s + 8 * lq);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(w + (16 * t + lr) * LD + 32 * s + 8 * lq);
            acc[t] = mfma(b, a, acc[t]);
        }
    }
}

// bias + ReLU, BF16, into the per-wave activation tile. The products come out transposed
// (tile_mfma_t: lane holds columns 16 t + 4 lq + j of row lr), so each lane writes 4
// adjacent columns of its row as one 8-byte store (a lane of the untransposed layout
// holds one column of 4 rows: 16 two-byte stores, two lanes per bank word).
__device__ __forceinline__ void store_hidden(__bf16 *act, const f32x4 *acc, const float *bias, int lr, int lq) {
#pragma unroll
    for (int t = 0; t < kGuideH / 16; ++t) {
        bf16x4 o;
#pragma unroll
        for (int j = 0; j < 4; ++j) o[j] = (__bf16)fmaxf(acc[t][j] + bias[16 * t + 4 * lq + j], 0.0f);
        *reinterpret_cast<bf16x4 *>(act + lr * kLdH + 16 * t + 4 * lq) = o;
    }
}

// orders a wave's LDS writes before its later reads (the tile is private to the wave)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(InferArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __bf16 *w1 = reinterpret_cast<__bf16 *>(lds_raw);
    __bf16 *w2 = w1 + kGuideH * kLd1;
    __bf16 *w3 = w2 + kGuideH * kLdH;
    __bf16 *w4 = w3 + kGuideH * kLdH;
    __bf16 *tiles = w4 + kGuideNOutPad16 * kLdH;
    float *b1 = reinterpret_cast<float *>(tiles + kInferWaves * kInferTile);
    float *b2 = b1 + kGuideH, *b3 = b2 + kGuideH, *b4 = b3 + kGuideH;
    stage_weights<kGuideK1, kLd1>(w1, a.w[0], kGuideH);
    stage_weights<kGuideH, kLdH>(w2, a.w[1], kGuideH);
    stage_weights<kGuideH, kLdH>(w3, a.w[2], kGuideH);
    stage_weights<kGuideH, kLdH>(w4, a.w[3], kGuideNOutPad16);
    for (int t = threadIdx.x; t < kGuideH; t += blockDim.x) {
        b1[t] = a.bias[0][t];
        b2[t] = a.bias[1][t];
        b3[t] = a.bias[2][t];
    }
    for (int t = threadIdx.x; t < kGuideNOutPad16; t += blockDim.x) b4[t] = a.bias[3][t];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    __bf16 *ft = tiles + wave * kInferTile;
    const int64_t rows = a.n_dev ? min((int64_t)*a.n_dev, a.n) : a.n;
    const int64_t ntiles = (rows + 15) >> 4;
    const int64_t stride = (int64_t)gridDim.x * kInferWaves;
    // A tile's inputs are loaded one tile ahead: the lane's 2 or 3 of the tile's 16 x 9
    // inputs (chain q = lane + 64 k = (row rr, input c), row-major, so the loads
    // coalesce), their rows' slots shuffled from the lanes that hold them (lane & 15 holds
    // row lane & 15). The next tile's slots load before this tile's first product and its
    // inputs after it, so neither the list nor the inputs stall the wave. Slot -1: a row
    // past the end (zero features, no output).
    auto row_slot = [&](int64_t tile) -> int32_t {
        const int64_t r = tile * 16 + lr;
        if (tile >= ntiles || r >= rows) return -1;
        return a.slot_of ? a.slot_of[r] : (int32_t)r;
    };
    float xin[3];
    auto load_inputs = [&](int32_t slot) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int q = lane + 64 * k, rr = q / kGuideIn, c = q - rr * kGuideIn;
            const int32_t src = __shfl(slot, rr & 15);
            xin[k] = (q < 16 * kGuideIn && src >= 0) ? a.in[(int64_t)src * kGuideIn + c] : 0.0f;
        }
    };
    int64_t tile = (int64_t)blockIdx.x * kInferWaves + wave;
    int32_t slot = row_slot(tile);
    load_inputs(slot);
    // whole waves only ever touch their own tile, so no block barrier is needed
    // inside the loop; wave_sync orders the tile's LDS writes and reads
    for (; tile < ntiles; tile += stride) {
        const int32_t slot_next = row_slot(tile + stride);
        // features x[rr][24 c + 2 f + {sin, cos}]; rows past the end get zero features
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int q = lane + 64 * k, rr = q / kGuideIn, c = q - rr * kGuideIn;
            const int32_t src = __shfl(slot, rr & 15);   // every lane takes part in the shuffle
            if (q < 16 * kGuideIn) {
                bf16x2 *o = reinterpret_cast<bf16x2 *>(ft + rr * kLd1 + 2 * kGuideFreq * c);
                if (src >= 0) {
                    FreqState st = freq_base(xin[k]);
#pragma unroll
                    for (int f = 0; f < kGuideFreq; ++f) {
                        o[f] = bf16x2{(__bf16)st.s, (__bf16)st.c};
                        freq_next(st);
                    }
                } else {
#pragma unroll
                    for (int f = 0; f < kGuideFreq; ++f) o[f] = bf16x2{(__bf16)0.0f, (__bf16)0.0f};
                }
            }
        }
        if (lane < 16) {
            bf16x8 z;
#pragma unroll
            for (int j = 0; j < 8; ++j) z[j] = (__bf16)0.0f;
            *reinterpret_cast<bf16x8 *>(ft + lane * kLd1 + kGuideFeat) = z;   // features 216..223
        }
        wave_sync();
        f32x4 acc[kGuideH / 16];
        tile_mfma_t<kGuideH / 16, kGuideK1, kLd1>(ft, w1, lr, lq, acc);
        load_inputs(slot_next);
        wave_sync();
        store_hidden(ft, acc, b1, lr, lq);
        wave_sync();
        if (a.h_out[0]) tile_to_global<kGuideH, kLdH>(ft, a.h_out[0], tile, lane);
        tile_mfma_t<kGuideH / 16, kGuideH, kLdH>(ft, w2, lr, lq, acc);
        wave_sync();
        store_hidden(ft, acc, b2, lr, lq);
        wave_sync();
        if (a.h_out[1]) tile_to_global<kGuideH, kLdH>(ft, a.h_out[1], tile, lane);
        tile_mfma_t<kGuideH / 16, kGuideH, kLdH>(ft, w3, lr, lq, acc);
        wave_sync();
        store_hidden(ft, acc, b3, lr, lq);
        wave_sync();
        if (a.h_out[2]) tile_to_global<kGuideH, kLdH>(ft, a.h_out[2], tile, lane);
        f32x4 out[kGuideNOutPad16 / 16];
        tile_mfma_t<kGuideNOutPad16 / 16, kGuideH, kLdH>(ft, w4, lr, lq, out);
        wave_sync();
        // the lane's row of raw: its slot, or, without a slot list, the tile row itself
        // also past the end (the buffers hold whole tiles)
        const int64_t dst = a.slot_of ? (int64_t)slot : tile * 16 + lr;
        if (dst >= 0) {
#pragma unroll
            for (int t = 0; t < kGuideNOutPad16 / 16; ++t)
                *reinterpret_cast<float4 *>(a.raw + dst * kGuideNOutPad16 + 16 * t + 4 * lq) =
                    make_float4(out[t][0] + b4[16 * t + 4 * lq], out[t][1] + b4[16 * t + 4 * lq + 1],
                                out[t][2] + b4[16 * t + 4 * lq + 2], out[t][3] + b4[16 * t + 4 * lq + 3]);
        }
        slot = slot_next;
    }
}

// loss = sum_r s_r * (-log q(w_r)) / sum_r s_r. Writes s_r d(-log q(w_r)) / d raw (bf16, 48
// wide + pad to 64), the row's weighted gradient before the division by sum_r s_r, which
// the reduction of the weight gradients applies, and per block the sums of
// s_r (-log q(w_r)) and of s_r, which k_loss_parts adds.
// With pb (the BSDF density of each row's direction), output kGuideAlphaOut also gets
// the gradient of sum_r s_r * (-log(alpha q + (1 - alpha) p_b)) / sum_r s_r, the
// cross-entropy of the one-sample MIS density, through alpha = learned_alpha(z): the
// selection probability falls where the mixture describes the incident radiance worse
// than the BSDF density does (Mueller et al. 2019, "Neural Importance Sampling").
// One thread per row; the kernel is bound by its VALU work, so exp, log, square root
// and reciprocal use the hardware approximations (about 1e-6 relative error, far below
// the BF16 rounding of the gradient), the log of each lobe's density is taken in the
// form log w + log kappa - log(2 pi (1 - e^{-2 kappa})) + kappa (mu.w - 1) with log kappa
// the clamped output itself, and only expm1 (in d/d kappa of the normaliser, where
// 1/kappa - 2/(e^{2 kappa} - 1) cancels for small kappa) stays the accurate function.
// A block's 256 rows of outputs are loaded into LDS with coalesced 16-byte loads and its
// 256 gradient rows leave the same way (a thread reading or writing its own row alone
// touches a different cache line in every lane).
constexpr int kLossRows = 256;
constexpr int kLossLdRaw = kGuideNOutPad16 + 1;   // floats per staged output row (odd: no bank conflicts)
constexpr int kLossLdGrad = kGuideNOutPad + 8;     // bf16 per staged gradient row (16-byte rows)
static_assert(kLossRows * kLossLdGrad * 2 <= kLossRows * kLossLdRaw * 4, "gradient rows reuse the output rows' LDS");
__global__ __launch_bounds__(kLossRows) void k_loss_grad(const float *__restrict__ raw, const float *__restrict__ dirs,
                                                   const float *__restrict__ wts, const float *__restrict__ pb,
                                                   __bf16 *__restrict__ draw, float *__restrict__ loss_part,
                                                   float *__restrict__ w_part, int64_t n) {
    const int64_t r0 = (int64_t)blockIdx.x * kLossRows;
    const int64_t r = r0 + threadIdx.x;
    __shared__ float red[kLossRows], redw[kLossRows];
    __shared__ __align__(16) float stage[kLossRows * kLossLdRaw];
    float contrib = 0.0f, wr = 0.0f;
    const int64_t rows_pad = (n + 15) & ~15LL;
    {
        constexpr int V = kGuideNOutPad16 / 4;   // 16-byte pieces per output row
        const float4 *src = reinterpret_cast<const float4 *>(raw + r0 * kGuideNOutPad16);
        for (int f = threadIdx.x; f < kLossRows * V; f += kLossRows) {
            const int row = f / V, c = 4 * (f % V);
            const float4 x = r0 + row < n ? src[f] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float *d = stage + row * kLossLdRaw + c;
            d[0] = x.x;
            d[1] = x.y;
            d[2] = x.z;
            d[3] = x.w;
        }
    }
    __syncthreads();
    // the row's 64 gradient columns (zero where nothing is written) go out as 8 x 16 B
    __bf16 gv[kGuideNOutPad];
#pragma unroll
    for (int c = 0; c < kGuideNOutPad; ++c) gv[c] = (__bf16)0.0f;
    if (r < n) {
        const float *rr = stage + threadIdx.x * kLossLdRaw;
        const float wx = dirs[3 * r], wy = dirs[3 * r + 1], wz = dirs[3 * r + 2];
        const float scale = wts[r];
        wr = scale;
        // mixture weights: softmax of the logits
        float mxl = -1e30f;
#pragma unroll
        for (int k = 0; k < kGuideLobes; ++k) mxl = fmaxf(mxl, rr[kGuideLobes * 4 + k]);
        float wk[kGuideLobes], wsum_k = 0.0f;
#pragma unroll
        for (int k = 0; k < kGuideLobes; ++k) {
            wk[k] = __expf(rr[kGuideLobes * 4 + k] - mxl);
            wsum_k += wk[k];
        }
        const float lws = __logf(wsum_k), iws = __builtin_amdgcn_rcpf(wsum_k);
        float lv[kGuideLobes], kap[kGuideLobes], ca[kGuideLobes], il[kGuideLobes], mu[kGuideLobes][3];
        float mx = -1e30f;
#pragma unroll
        for (int k = 0; k < kGuideLobes; ++k) {
            const float m0 = rr[3 * k], m1 = rr[3 * k + 1], m2 = rr[3 * k + 2];
            il[k] = __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(m0 * m0 + m1 * m1 + m2 * m2) + 1e-6f);
            mu[k][0] = m0 * il[k];
            mu[k][1] = m1 * il[k];
            mu[k][2] = m2 * il[k];
            const float lk = fminf(fmaxf(rr[kGuideLobes * 3 + k], kGuideLogKappaMin), kGuideLogKappaMax);
            kap[k] = __expf(lk);
            ca[k] = mu[k][0] * wx + mu[k][1] * wy + mu[k][2] * wz;
            lv[k] = (rr[kGuideLobes * 4 + k] - mxl - lws) + lk - 1.83787706641f -
                    __logf(1.0f - __expf(-2.0f * kap[k])) + kap[k] * (ca[k] - 1.0f);
            mx = fmaxf(mx, lv[k]);
        }
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < kGuideLobes; ++k) s += __expf(lv[k] - mx);
        const float logq = mx + __logf(s);
        contrib = scale * (-logq);
        const float lk_min = kGuideLogKappaMin, lk_max = kGuideLogKappaMax;
#pragma unroll
        for (int k = 0; k < kGuideLobes; ++k) {
            const float resp = __expf(lv[k] - logq);
            // logits: d(-log q)/d logit_k = w_k - resp_k
            gv[kGuideLobes * 4 + k] = (__bf16)(scale * (wk[k] * iws - resp));
            // log kappa: d log v / d kappa = 1/k - 2/(e^{2k} - 1) + (mu.w - 1); d kappa / d lk = kappa
            const float dlv = __builtin_amdgcn_rcpf(kap[k]) - 2.0f / expm1f(2.0f * kap[k]) + (ca[k] - 1.0f);
            const float lk = rr[kGuideLobes * 3 + k];
            const float gk = (lk > lk_min && lk < lk_max) ? -scale * resp * dlv * kap[k] : 0.0f;
            gv[kGuideLobes * 3 + k] = (__bf16)gk;
            // raw mean m: d log v / d m = kappa (w - (mu.w) mu) / |m|
            const float f = -scale * resp * kap[k] * il[k];
            gv[3 * k] = (__bf16)(f * (wx - ca[k] * mu[k][0]));
            gv[3 * k + 1] = (__bf16)(f * (wy - ca[k] * mu[k][1]));
            gv[3 * k + 2] = (__bf16)(f * (wz - ca[k] * mu[k][2]));
        }
        if (pb) {
            const float q = __expf(logq), p_b = pb[r];
            const float sg = sigmoidf_(rr[kGuideAlphaOut]);
            const float al = kAlphaLo + (kAlphaHi - kAlphaLo) * sg;
            const float pc = fmaxf(al * q + (1.0f - al) * p_b, 1e-30f);
            const float ga = -scale * (q - p_b) / pc * (kAlphaHi - kAlphaLo) * sg * (1.0f - sg);
            gv[kGuideAlphaOut] = (__bf16)(isfinite(ga) ? ga : 0.0f);
        }
    }
    __syncthreads();   // every row's outputs are read: the staging area takes the gradients
    __bf16 *gs = reinterpret_cast<__bf16 *>(stage);
#pragma unroll
    for (int v = 0; v < kGuideNOutPad / 8; ++v) {
        bf16x8 w;
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = gv[8 * v + j];
        *reinterpret_cast<bf16x8 *>(gs + threadIdx.x * kLossLdGrad + 8 * v) = w;
    }
    __syncthreads();
    {
        constexpr int V = kGuideNOutPad / 8;   // 16-byte pieces per gradient row
        bf16x8 *dst = reinterpret_cast<bf16x8 *>(draw + r0 * kGuideNOutPad);
        for (int f = threadIdx.x; f < kLossRows * V; f += kLossRows) {
            const int row = f / V, c = 8 * (f % V);
            if (r0 + row < rows_pad) dst[f] = *reinterpret_cast<const bf16x8 *>(gs + row * kLossLdGrad + c);
        }
    }
    red[threadIdx.x] = contrib;
    redw[threadIdx.x] = wr;
    __syncthreads();
    for (int o = kLossRows / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            red[threadIdx.x] += red[threadIdx.x + o];
            redw[threadIdx.x] += redw[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        loss_part[blockIdx.x] = red[0];
        w_part[blockIdx.x] = redw[0];
    }
}

// The loss and the weight sum from k_loss_grad's nb block partials, added in a fixed
// order by one block (no float atomics), so both are the same bits on every run:
// scalars[1] = sum_r s_r, scalars[0] = sum_r s_r (-log q(w_r)) / scalars[1].
__global__ __launch_bounds__(256) void k_loss_parts(const float *__restrict__ loss_part,
                                                    const float *__restrict__ w_part, int64_t nb,
                                                    float *__restrict__ scalars) {
    __shared__ float red[256], redw[256];
    float sl = 0.0f, sw = 0.0f;
    for (int64_t i = threadIdx.x; i < nb; i += 256) {
        sl += loss_part[i];
        sw += w_part[i];
    }
    red[threadIdx.x] = sl;
    redw[threadIdx.x] = sw;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            red[threadIdx.x] += red[threadIdx.x + o];
            redw[threadIdx.x] += redw[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scalars[1] = redw[0];
        scalars[0] = red[0] / fmaxf(redw[0], 1e-30f);
    }
}

// ------------------------------------------------------------------ weight gradients
// partial[blk][i][o] = sum over the block's rows r of H[r][i] G[r][o]  (KI x KO) and
// pbias[blk][o] = sum over the block's rows of G[r][o]. Rows are staged kWgRows at
// a time in LDS, transposed ([feature][row]) so that each lane's 8 operand values
// along the row (reduction) dimension are 16 contiguous bytes. A staging task moves
// 8 rows x 8 features: eight 16-byte loads into registers, an 8 x 8 transpose in
// registers, eight 16-byte LDS rows. Consecutive threads take consecutive row groups
// of one feature group, so the 16 lanes of a quarter wave store to 16 different LDS
// bank groups (with the feature groups fastest, 4 lanes shared each group). The tasks of
// H and G are spread over all of the block's threads, and the next chunk's loads are issued before the current chunk's
// products, so that they are in flight while the matrix cores work. The bias sums
// are products too: an all-ones A operand against a tile of G gives G's column sums
// in every row of the result.
constexpr int kWgRows = 128;
constexpr int kWgThreads = 256;

template <int KI, int KO, bool ENC>
struct WgStage {
    static constexpr int RG = kWgRows / 8;                 // row groups of a chunk
    static constexpr int TH = ENC ? 0 : RG * (KI / 8);     // H tasks (none when the features are computed)
    static constexpr int TG = RG * (KO / 8);               // G tasks
    static constexpr int NT = (TH + TG + kWgThreads - 1) / kWgThreads;   // tasks per thread
};

// Rows 8 rg .. 8 rg + 7 of the chunk at rb, features 8 fg .. 8 fg + 7 (zero at or past r1).
__device__ __forceinline__ void wg_load(bf16x8 (&v)[8], const __bf16 *__restrict__ src, int ld, int rg, int fg,
                                        int64_t rb, int64_t r1) {
#pragma unroll
    for (int rr = 0; rr < 8; ++rr) {
        const int64_t row = rb + 8 * rg + rr;
        if (row < r1) {
            v[rr] = *reinterpret_cast<const bf16x8 *>(src + row * ld + 8 * fg);
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) v[rr][q] = (__bf16)0.0f;
        }
    }
}

template <int LD>
__device__ __forceinline__ void wg_store(__bf16 (*dst)[LD], const bf16x8 (&v)[8], int rg, int fg) {
#pragma unroll
    for (int ff = 0; ff < 8; ++ff) {
        bf16x8 w;
#pragma unroll
        for (int rr = 0; rr < 8; ++rr) w[rr] = v[rr][ff];
        *reinterpret_cast<bf16x8 *>(&dst[8 * fg + ff][8 * rg]) = w;
    }
}

// The first layer's input features of rows [rb, rb + kWgRows) straight from the
// inputs, [feature][row] in LDS: the encoding of k_encode / k_mlp_infer (the same
// bits), zero for rows at or past `valid` (the batch's rows, or the block's end). A
// task is one input of two adjacent rows, rows fastest across the threads, and writes
// each feature of the pair as one 4-byte word. The padding features are not written.
__device__ __forceinline__ void stage_features(__bf16 (*dst)[kWgRows + 8], const float *__restrict__ in, int64_t rb,
                                               int64_t valid) {
    constexpr int PAIRS = kWgRows / 2;
    for (int task = threadIdx.x; task < PAIRS * kGuideIn; task += blockDim.x) {
        const int rp = task % PAIRS, c = task / PAIRS;
        const int64_t row = rb + 2 * rp;
        const bool ok0 = row < valid, ok1 = row + 1 < valid;
        FreqState a = freq_base(ok0 ? in[row * kGuideIn + c] : 0.0f);
        FreqState b = freq_base(ok1 ? in[(row + 1) * kGuideIn + c] : 0.0f);
        for (int f = 0; f < kGuideFreq; ++f) {
            const int k = 2 * kGuideFreq * c + 2 * f;
            bf16x2 sv, cv;
            sv[0] = (__bf16)(ok0 ? a.s : 0.0f);
            sv[1] = (__bf16)(ok1 ? b.s : 0.0f);
            cv[0] = (__bf16)(ok0 ? a.c : 0.0f);
            cv[1] = (__bf16)(ok1 ? b.c : 0.0f);
            *reinterpret_cast<bf16x2 *>(&dst[k][2 * rp]) = sv;
            *reinterpret_cast<bf16x2 *>(&dst[k + 1][2 * rp]) = cv;
            freq_next(a);
            freq_next(b);
        }
    }
}

// ENC (first layer only): H is not read; the features come from enc_in (rows below
// enc_rows) through stage_features.
template <int KI, int KO, bool ENC = false>
__global__ __launch_bounds__(kWgThreads) void k_wgrad(const __bf16 *__restrict__ H, int ldh,
                                                      const __bf16 *__restrict__ G, int ldg, int64_t n,
                                                      int64_t rows_per_block, float *__restrict__ partial,
                                                      float *__restrict__ pbias, const float *__restrict__ enc_in = nullptr,
                                                      int64_t enc_rows = 0) {
    using S = WgStage<KI, KO, ENC>;
    constexpr int TI = KI / 16, TO = KO / 16, TILES = TI * TO;
    constexpr int PER_WAVE = (TILES + TO + 3) / 4;   // the weight tiles, then one bias tile per output tile
    __shared__ __align__(16) __bf16 hs[KI][kWgRows + 8];  // [feature][row]
    __shared__ __align__(16) __bf16 gs[KO][kWgRows + 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r1 = min(n, r0 + rows_per_block);
    if (ENC) {
        // the padding features stay zero for every chunk
        for (int i = threadIdx.x; i < (KI - kGuideFeat) * (kWgRows + 8); i += kWgThreads)
            hs[kGuideFeat + i / (kWgRows + 8)][i % (kWgRows + 8)] = (__bf16)0.0f;
    }
    bf16x8 ones;
#pragma unroll
    for (int q = 0; q < 8; ++q) ones[q] = (__bf16)1.0f;
    f32x4 acc[PER_WAVE];
#pragma unroll
    for (int q = 0; q < PER_WAVE; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};

    bf16x8 v[S::NT][8];
    // the thread's staging tasks of the chunk at rb, into registers
#define TT_WG_LOAD(rb_)                                                                            \
    _Pragma("unroll") for (int k = 0; k < S::NT; ++k) {                                            \
        const int t = threadIdx.x + k * kWgThreads;                                                \
        if (t < S::TH) wg_load(v[k], H, ldh, t % S::RG, t / S::RG, (rb_), r1);                     \
        else if (t < S::TH + S::TG)                                                                \
            wg_load(v[k], G, ldg, (t - S::TH) % S::RG, (t - S::TH) / S::RG, (rb_), r1);             \
    }
    if (r0 < r1) TT_WG_LOAD(r0)
    for (int64_t rb = r0; rb < r1; rb += kWgRows) {
        __syncthreads();   // the previous chunk's products have read hs and gs
#pragma unroll
        for (int k = 0; k < S::NT; ++k) {
            const int t = threadIdx.x + k * kWgThreads;
            if (t < S::TH) wg_store(hs, v[k], t % S::RG, t / S::RG);
            else if (t < S::TH + S::TG) wg_store(gs, v[k], (t - S::TH) % S::RG, (t - S::TH) / S::RG);
        }
        if (ENC) stage_features(hs, enc_in, rb, min(r1, enc_rows));
        __syncthreads();
        if (rb + kWgRows < r1) TT_WG_LOAD(rb + kWgRows)
#pragma unroll
        for (int q = 0; q < PER_WAVE; ++q) {
            const int tile = wave * PER_WAVE + q;
            if (tile < TILES) {
                const int ti = tile / TO, to = tile % TO;
#pragma unroll
                for (int s2 = 0; s2 < kWgRows / 32; ++s2) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&hs[16 * ti + lr][32 * s2 + 8 * lq]);
                    const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&gs[16 * to + lr][32 * s2 + 8 * lq]);
                    acc[q] = mfma(a, b, acc[q]);
                }
            } else if (tile < TILES + TO) {
                const int to = tile - TILES;
#pragma unroll
                for (int s2 = 0; s2 < kWgRows / 32; ++s2) {
                    const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&gs[16 * to + lr][32 * s2 + 8 * lq]);
                    acc[q] = mfma(ones, b, acc[q]);
                }
            }
        }
    }
#undef TT_WG_LOAD
    float *out = partial + (int64_t)blockIdx.x * KI * KO;
#pragma unroll
    for (int q = 0; q < PER_WAVE; ++q) {
        const int tile = wave * PER_WAVE + q;
        if (tile < TILES) {
            const int ti = tile / TO, to = tile % TO;
#pragma unroll
            for (int j = 0; j < 4; ++j) out[(16 * ti + 4 * lq + j) * KO + 16 * to + lr] = acc[q][j];
        } else if (tile < TILES + TO && lq == 0) {
            pbias[(int64_t)blockIdx.x * KO + 16 * (tile - TILES) + lr] = acc[q][0];
        }
    }
}

// ------------------------------------------------------------------ fused backward
// The data gradients of layers 4, 3, 2 and the weight gradients of layers 4, 3, 2 in one
// kernel. For every ROWS-row chunk of the block's rows:
//   1. H3, H2, H1 and dRaw of the chunk are staged transposed ([feature][row]) in LDS, as
//      k_wgrad stages its operands (the next chunk's loads are issued before the products);
//   2. each wave takes one 16-row tile through the three backward products
//      dH3 = (dRaw W4) * (H3 > 0), dH2 = (dH3 W3) * (H2 > 0), dH1 = (dH2 W2) * (H1 > 0): each
//      product's output kept row-major in the wave's tile as the next product's A operand
//      and written transposed into LDS (dH3, dH2) for the weight gradients; the ReLU masks
//      come from the staged activations; dH1 goes to global, where the first layer's
//      weight gradient reads it. The backward weight copies are read from LDS (WLDS) or
//      straight from global memory, where all waves of the chip share them in cache;
//   3. the weight gradients dW4 += H3^T dRaw, dW3 += H2^T dH3, dW2 += H1^T dH2 and their
//      bias sums accumulate over the chunk, the 55 result tiles spread over the waves.
// The per-tile arithmetic is that of k_gemm; each result tile takes its 32-row MFMA slices
// in ascending row order over the block's rows, as k_wgrad does (for any chunk size), so
// dH1 and the partials are the same bits as the separate kernels' (the unfused path
// computes them that way), and dH3, dH2 never leave the CU.
constexpr int kBwTiles = 3 * 4 + 4 * 4 + 4 * 4 + 3 + 4 + 4;   // weight tiles of layers 4, 3, 2, then bias tiles

template <int ROWS, bool WLDS>
struct BwShape {
    static constexpr int WAVES = ROWS / 16;   // one 16-row tile of the chunk per wave
    static constexpr int LDT = ROWS + 8;      // transposed staging row stride (bf16)
    static constexpr int PER_WAVE = (kBwTiles + WAVES - 1) / WAVES;
    static constexpr int LDW = WLDS ? kLdH : kGuideH;   // row stride of the backward weight copies read
    static constexpr int LDS =
        2 * ((WLDS ? 3 * kGuideH * kLdH : 0) + (5 * kGuideH + kGuideNOutPad16) * LDT + WAVES * 16 * kLdH);
    static_assert(LDS <= 160 * 1024, "fused backward block exceeds the CU's LDS");
    static_assert(ROWS % 32 == 0 && kWgRows % ROWS == 0, "whole MFMA row slices; chunks tile the block's rows");
};

struct BwdArgs {
    const __bf16 *draw;      // [R][64] output gradients
    const __bf16 *wb[3];     // backward copies of layers 2, 3, 4: [in 64][out_pad 64]
    const __bf16 *h[3];      // H1, H2, H3 [R][64]
    __bf16 *dh1;             // dH1 [R][64]
    int64_t rows;            // R, a multiple of 16
    int64_t rows_per_block;  // the weight gradients' block rows (a multiple of kWgRows)
    float *pw[3], *pb[3];    // weight / bias partials of layers 2, 3, 4 (block b at + b * count)
};

// One backward product's output (D layout: rows 4 lq + j, column 16 t + lr of the wave's
// tile, chunk row `cr` for j = 0): rounded to BF16, masked where the layer's activation
// is not positive, into the wave's row-major tile and, if td is set, transposed into td.
template <int LDT>
__device__ __forceinline__ void bwd_out(const f32x4 *acc, const __bf16 (*th)[LDT], __bf16 (*td)[LDT], __bf16 *ft,
                                        int cr, int lr, int lq) {
#pragma unroll
    for (int t = 0; t < kGuideH / 16; ++t) {
        const int col = 16 * t + lr;
        const bf16x4 hm = *reinterpret_cast<const bf16x4 *>(&th[col][cr]);
        bf16x4 o;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            __bf16 v = (__bf16)(acc[t][j] + 0.0f);
            if (!((float)hm[j] > 0.0f)) v = (__bf16)0.0f;
            o[j] = v;
            ft[(4 * lq + j) * kLdH + col] = v;
        }
        if (td) *reinterpret_cast<bf16x4 *>(&td[col][cr]) = o;
    }
}

// The wave's 16-row tile (row-major in LDS, stride kLdH) times W^T over K = 64, W rows
// of stride LDW (LDS or global)
template <int LDW>
__device__ __forceinline__ void tile_mfma_w(const __bf16 *act, const __bf16 *w, int lr, int lq, f32x4 *acc) {
#pragma unroll
    for (int t = 0; t < kGuideH / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < kGuideH / 32; ++s) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(act + lr * kLdH + 32 * s + 8 * lq);
#pragma unroll
        for (int t = 0; t < kGuideH / 16; ++t) {
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(w + (16 * t + lr) * LDW + 32 * s + 8 * lq);
            acc[t] = mfma(a, b, acc[t]);
        }
    }
}

template <int ROWS, bool WLDS>
__global__ __launch_bounds__(ROWS * 4) void k_bwd_wgrad(BwdArgs a) {
    using S = BwShape<ROWS, WLDS>;
    constexpr int LDT = S::LDT, LDW = S::LDW;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __bf16 *wl = reinterpret_cast<__bf16 *>(lds_raw);
    const __bf16 *w2 = WLDS ? wl : a.wb[0];                          // layer 2 (dH2 -> dH1)
    const __bf16 *w3 = WLDS ? wl + kGuideH * kLdH : a.wb[1];         // layer 3 (dH3 -> dH2)
    const __bf16 *w4 = WLDS ? wl + 2 * kGuideH * kLdH : a.wb[2];     // layer 4 (dRaw -> dH3)
    typedef __bf16 TRow[LDT];
    TRow *tH3 = reinterpret_cast<TRow *>(wl + (WLDS ? 3 * kGuideH * kLdH : 0));
    TRow *tH2 = tH3 + kGuideH, *tH1 = tH2 + kGuideH, *tD3 = tH1 + kGuideH, *tD2 = tD3 + kGuideH;
    TRow *tD4 = tD2 + kGuideH;   // dRaw, the head's kGuideNOutPad16 columns
    __bf16 *tiles = reinterpret_cast<__bf16 *>(tD4 + kGuideNOutPad16);
    if (WLDS) {
        stage_weights<kGuideH, kLdH>(wl, a.wb[0], kGuideH);
        stage_weights<kGuideH, kLdH>(wl + kGuideH * kLdH, a.wb[1], kGuideH);
        stage_weights<kGuideNOutPad, kLdH>(wl + 2 * kGuideH * kLdH, a.wb[2], kGuideH);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    __bf16 *ft = tiles + wave * 16 * kLdH;
    const int64_t r0 = (int64_t)blockIdx.x * a.rows_per_block;
    const int64_t r1 = min(a.rows, r0 + a.rows_per_block);
    bf16x8 ones;
#pragma unroll
    for (int q = 0; q < 8; ++q) ones[q] = (__bf16)1.0f;
    f32x4 acc[S::PER_WAVE];
#pragma unroll
    for (int q = 0; q < S::PER_WAVE; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};

    // staging tasks (8 rows x 8 features each): H3, H2, H1 (8 feature groups), dRaw (6)
    constexpr int RG = ROWS / 8, FH = kGuideH / 8, FD = kGuideNOutPad16 / 8;
    constexpr int TASKS = RG * (3 * FH + FD);
    static_assert(TASKS <= 64 * S::WAVES, "one staging task per thread");
    const int task = threadIdx.x;
    const int which = task < RG * FH ? 0 : task < 2 * RG * FH ? 1 : task < 3 * RG * FH ? 2 : task < TASKS ? 3 : 4;
    const int tt = which < 3 ? task - which * RG * FH : task - 3 * RG * FH;
    const int rg = tt % RG, fg = tt / RG;
    const __bf16 *src = which == 0 ? a.h[2] : which == 1 ? a.h[1] : which == 2 ? a.h[0] : a.draw;
    const int ld = which < 3 ? kGuideH : kGuideNOutPad;
    TRow *dst = which == 0 ? tH3 : which == 1 ? tH2 : which == 2 ? tH1 : tD4;
    bf16x8 v[8];
    if (which < 4 && r0 < r1) wg_load(v, src, ld, rg, fg, r0, r1);

    for (int64_t rb = r0; rb < r1; rb += ROWS) {
        __syncthreads();   // the weights are staged; the previous chunk's products are done with LDS
        if (which < 4) wg_store(dst, v, rg, fg);
        __syncthreads();
        if (which < 4 && rb + ROWS < r1) wg_load(v, src, ld, rg, fg, rb + ROWS, r1);

        // 2. the wave's tile through the three backward products
        const int cr = 16 * wave + 4 * lq;   // chunk row of the lane's first output row
        const int64_t trow = rb + 16 * wave;
        if (trow < r1) {
            f32x4 d[kGuideH / 16];
#pragma unroll
            for (int t = 0; t < kGuideH / 16; ++t) d[t] = f32x4{0.f, 0.f, 0.f, 0.f};
            const __bf16 *arow = a.draw + (trow + lr) * kGuideNOutPad + 8 * lq;
#pragma unroll
            for (int s2 = 0; s2 < kGuideNOutPad / 32; ++s2) {
                const bf16x8 av = *reinterpret_cast<const bf16x8 *>(arow + 32 * s2);
#pragma unroll
                for (int t = 0; t < kGuideH / 16; ++t) {
                    const bf16x8 b = *reinterpret_cast<const bf16x8 *>(w4 + (16 * t + lr) * LDW + 32 * s2 + 8 * lq);
                    d[t] = mfma(av, b, d[t]);
                }
            }
            bwd_out<LDT>(d, tH3, tD3, ft, cr, lr, lq);
            wave_sync();
            tile_mfma_w<LDW>(ft, w3, lr, lq, d);
            wave_sync();
            bwd_out<LDT>(d, tH2, tD2, ft, cr, lr, lq);
            wave_sync();
            tile_mfma_w<LDW>(ft, w2, lr, lq, d);
            wave_sync();
            bwd_out<LDT>(d, tH1, nullptr, ft, cr, lr, lq);
            wave_sync();
            tile_to_global<kGuideH, kLdH>(ft, a.dh1, trow >> 4, lane);
        } else {
            // rows past the batch: zero gradients for the weight products
            bf16x4 z;
#pragma unroll
            for (int j = 0; j < 4; ++j) z[j] = (__bf16)0.0f;
#pragma unroll
            for (int t = 0; t < kGuideH / 16; ++t) {
                *reinterpret_cast<bf16x4 *>(&tD3[16 * t + lr][cr]) = z;
                *reinterpret_cast<bf16x4 *>(&tD2[16 * t + lr][cr]) = z;
            }
        }
        __syncthreads();   // dH3, dH2 of every tile are in LDS

        // 3. weight gradients over the chunk: tiles 0..11 layer 4 (4 x 3), 12..27 layer 3,
        // 28..43 layer 2 (4 x 4 each), then bias tiles 44..46, 47..50, 51..54
#pragma unroll
        for (int q = 0; q < S::PER_WAVE; ++q) {
            const int tile = wave * S::PER_WAVE + q;
            if (tile >= kBwTiles) continue;
            const TRow *A = nullptr, *B;
            int ti = 0, to;
            if (tile < 12) {
                A = tH3, B = tD4, ti = tile / 3, to = tile % 3;
            } else if (tile < 28) {
                A = tH2, B = tD3, ti = (tile - 12) / 4, to = (tile - 12) % 4;
            } else if (tile < 44) {
                A = tH1, B = tD2, ti = (tile - 28) / 4, to = (tile - 28) % 4;
            } else if (tile < 47) {
                B = tD4, to = tile - 44;
            } else if (tile < 51) {
                B = tD3, to = tile - 47;
            } else {
                B = tD2, to = tile - 51;
            }
#pragma unroll
            for (int s2 = 0; s2 < ROWS / 32; ++s2) {
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&B[16 * to + lr][32 * s2 + 8 * lq]);
                const bf16x8 av = A ? *reinterpret_cast<const bf16x8 *>(&A[16 * ti + lr][32 * s2 + 8 * lq]) : ones;
                acc[q] = mfma(av, b, acc[q]);
            }
        }
    }

    // partials in k_wgrad's layout: weights [in][out] per block, biases [out] per block
#pragma unroll
    for (int q = 0; q < S::PER_WAVE; ++q) {
        const int tile = wave * S::PER_WAVE + q;
        if (tile >= kBwTiles) continue;
        if (tile < 44) {
            const int layer = tile < 12 ? 2 : tile < 28 ? 1 : 0;   // index into pw: layers 2, 3, 4
            const int ko = layer == 2 ? kGuideNOutPad16 : kGuideH;
            const int lt = tile < 12 ? tile : tile < 28 ? tile - 12 : tile - 28;
            const int ti = lt / (ko / 16), to = lt % (ko / 16);
            float *out = a.pw[layer] + (int64_t)blockIdx.x * kGuideH * ko;
#pragma unroll
            for (int j = 0; j < 4; ++j) out[(16 * ti + 4 * lq + j) * ko + 16 * to + lr] = acc[q][j];
        } else if (lq == 0) {
            const int layer = tile < 47 ? 2 : tile < 51 ? 1 : 0;
            const int ko = layer == 2 ? kGuideNOutPad16 : kGuideH;
            const int to = tile < 47 ? tile - 44 : tile < 51 ? tile - 47 : tile - 51;
            a.pb[layer][(int64_t)blockIdx.x * ko + 16 * to + lr] = acc[q][0];
        }
    }
}

// The fused backward's chunk rows and weight placement: 128-row chunks (8 waves) with
// the weights in LDS took 77 us per 2^18-row step; 64-row chunks 113 us (weights in LDS)
// and 136 us (from global), 32-row chunks 134 us, 128-row chunks with the weights read
// from global 101 us: every wave reads each weight matrix once per product, which the
// vector memory pipeline serves more slowly than LDS.
constexpr int kBwRows = 128;
constexpr bool kBwWLds = true;

// Sum of the nb per-block partial gradients of every parameter, in a fixed order
// (deterministic), all layers in one launch: 8 regions of partials (each layer's
// weights [in][out] and biases [out], region r at partial + start[r], block b's values
// at + b * count[r]). A block takes 64 elements of the concatenated regions, its
// kRedParts waves each sum every kRedParts-th partial (unrolled, so several loads are
// in flight per lane), and the first wave adds the part sums in part order, divides by
// the training weights' sum (*wsum: the rows' output gradients are weighted by s_r
// alone) and transposes weights to the [out][in] master layout.
constexpr int kRedParts = 16;
struct RedRegions {
    int64_t start[8];
    int count[8], first[8];   // elements, and the first element of the region in the concatenation
    int dst[8], ki[8], ko[8];
    int transpose[8];
};
__global__ __launch_bounds__(64 * kRedParts) void k_reduce_all(const float *__restrict__ partial, int nb, int total,
                                                               RedRegions R, const float *__restrict__ wsum,
                                                               float *__restrict__ grad) {
    __shared__ float red[kRedParts][64];
    const int el = threadIdx.x & 63, part = threadIdx.x >> 6;
    const int e = blockIdx.x * 64 + el;
    int r = 0;
#pragma unroll
    for (int q = 1; q < 8; ++q)
        if (e >= R.first[q]) r = q;
    const int le = e - R.first[r];
    float s = 0.0f;
    if (e < total) {
        const float *src = partial + R.start[r] + le;
        const int64_t stride = R.count[r];
#pragma unroll 8
        for (int b = part; b < nb; b += kRedParts) s += src[(int64_t)b * stride];
    }
    red[part][el] = s;
    __syncthreads();
    if (part == 0 && e < total) {
        float t = red[0][el];
#pragma unroll
        for (int q = 1; q < kRedParts; ++q) t += red[q][el];
        int d = le;
        if (R.transpose[r]) {
            const int i = le / R.ko[r], o = le % R.ko[r];
            d = o * R.ki[r] + i;
        }
        grad[R.dst[r] + d] = t / fmaxf(*wsum, 1e-30f);
    }
}

// ------------------------------------------------------------------ Adam + BF16 copies
// Adam step on every parameter, and the BF16 copies of the updated weights:
// forward Bt[out][in] and backward W[in][out_pad] of the layer the element is in.
struct AdamLayers {
    int off[4], in[4], out[4], out_pad_bwd[4];
    __bf16 *fwd[4], *bwd[4];
};

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                                              const float *__restrict__ g, int n, float lr, float b1, float b2,
                                              float eps, float c1, float c2, AdamLayers L) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.0f - b1) * gg;
    const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    const float np = p[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    p[e] = np;
    int k = 0;
#pragma unroll
    for (int q = 1; q < 4; ++q)
        if (e >= L.off[q]) k = q;
    const int le = e - L.off[k];
    if (le < L.in[k] * L.out[k]) {
        const int o = le / L.in[k], i = le % L.in[k];
        const __bf16 bv = (__bf16)np;
        L.fwd[k][le] = bv;
        L.bwd[k][i * L.out_pad_bwd[k] + o] = bv;
    }
}

// Master layout per layer: W[out][in] then b[out]. Forward copy Bt = W[out][in]
// (rows padded to the padded out width), backward copy W[in][out_pad].
__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ p, __bf16 *__restrict__ fwd,
                                                 __bf16 *__restrict__ bwd, int in, int out, int out_pad_bwd) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= in * out) return;
    const int o = e / in, i = e % in;
    const __bf16 b = (__bf16)p[e];
    fwd[e] = b;
    bwd[i * out_pad_bwd + o] = b;
}

}  // namespace

// ------------------------------------------------------------------ host side
const GuideLayer kGuideLayers[4] = {
    {kGuideK1, kGuideH}, {kGuideH, kGuideH}, {kGuideH, kGuideH}, {kGuideH, kGuideNOutPad16}};

int64_t guide_param_count() {
    int64_t n = 0;
    for (const GuideLayer &l : kGuideLayers) n += (int64_t)l.in * l.out + l.out;
    return n;
}

static int64_t layer_offset(int k) {
    int64_t n = 0;
    for (int q = 0; q < k; ++q) n += (int64_t)kGuideLayers[q].in * kGuideLayers[q].out + kGuideLayers[q].out;
    return n;
}

tt_status guide_alloc_batch(tt_guide *g, int64_t rows) {
    if (g->cap >= rows) return TT_OK;
    GuideBatch &b = g->batch;
    void *ptrs[] = {b.in,   b.x,     b.h[0],  b.h[1],  b.h[2], b.raw, b.draw,
                    b.dh[0], b.dh[1], b.dh[2], b.dirs, b.wts,  b.pb,  b.lpart};
    for (void *q : ptrs)
        if (q) hipFree(q);
    b = GuideBatch();
    g->cap = 0;
    const int64_t R = (rows + 15) & ~15LL;
    tt_status st = TT_OK;
    auto al = [&](void **p, size_t bytes, const char *w) {
        if (st == TT_OK) st = dev_alloc(g->ctx, p, bytes, w);
    };
    al((void **)&b.in, (size_t)R * kGuideIn * 4, "guide inputs");
    al((void **)&b.x, (size_t)R * kGuideK1 * 2, "guide features");
    for (int k = 0; k < 3; ++k) al((void **)&b.h[k], (size_t)R * kGuideH * 2, "guide activations");
    al((void **)&b.raw, (size_t)R * kGuideNOutPad16 * 4, "guide outputs");
    al((void **)&b.draw, (size_t)R * kGuideNOutPad * 2, "guide output gradients");
    for (int k = 0; k < 3; ++k) al((void **)&b.dh[k], (size_t)R * kGuideH * 2, "guide activation gradients");
    al((void **)&b.dirs, (size_t)R * 3 * 4, "guide directions");
    al((void **)&b.wts, (size_t)R * 4, "guide weights");
    al((void **)&b.pb, (size_t)R * 4, "guide BSDF densities");
    al((void **)&b.lpart, (size_t)2 * ((R + kLossRows - 1) / kLossRows) * 4, "guide loss partials");
    if (st == TT_OK) g->cap = R;
    return st;
}

// Inference rows of the renderer: about 232 B each (inputs, mixture, flag bit and slot
// index), so a render can keep as many paths in flight as the tracer itself; the
// training batch above (about 1.6 KB a row) is sized by the Adam batch alone. Rows are
// counted in whole 256-slot steps of the guide-row kernels (each step's 4 flag words are
// all written), which run at most 16 blocks of 256 threads per CU.
tt_status guide_alloc_infer(tt_guide *g, int64_t rows) {
    if (g->icap >= rows) return TT_OK;
    void *old[] = {g->iin, g->iraw, g->imask, g->iblock_n, g->islot};
    for (void *q : old)
        if (q) hipFree(q);
    g->iin = g->iraw = nullptr;
    g->imask = nullptr;
    g->iblock_n = g->islot = nullptr;
    g->icap = g->iblocks = 0;
    const int64_t R = (rows + 255) & ~255LL;
    if (R >= INT32_MAX) return fail(TT_ERR_INVALID, "guide inference: too many queue slots");
    const int64_t blocks = 16LL * std::max(1, g->ctx->num_cus);
    tt_status st = dev_alloc(g->ctx, (void **)&g->iin, (size_t)R * kGuideIn * 4, "guide inference inputs");
    if (st == TT_OK) st = dev_alloc(g->ctx, (void **)&g->iraw, (size_t)R * kGuideNOutPad16 * 4, "guide inference outputs");
    if (st == TT_OK) st = dev_alloc(g->ctx, (void **)&g->imask, (size_t)R / 8, "guide inference row flags");
    if (st == TT_OK) st = dev_alloc(g->ctx, (void **)&g->iblock_n, (size_t)(blocks + 1) * 4, "guide inference row counts");
    if (st == TT_OK) st = dev_alloc(g->ctx, (void **)&g->islot, (size_t)R * 4, "guide inference row slots");
    if (st == TT_OK) {
        g->icap = R;
        g->iblocks = blocks;
    }
    return st;
}

static hipError_t refresh_bf16(tt_guide *g, hipStream_t s) {
    for (int k = 0; k < 4; ++k) {
        const GuideLayer &l = kGuideLayers[k];
        const int n = l.in * l.out;
        const int out_pad_bwd = k == 3 ? kGuideNOutPad : l.out;
        hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, s, g->params + layer_offset(k), g->wf[k],
                           g->wb[k], l.in, l.out, out_pad_bwd);
    }
    return hipGetLastError();
}

// Forward over the first n rows of g->batch.in (device count n_dev optional):
// features, 3 hidden layers, raw outputs (fp32, 48 wide).
hipError_t guide_forward(tt_guide *g, int64_t n, const int32_t *n_dev, hipStream_t s) {
    GuideBatch &b = g->batch;
    const int64_t R = (n + 15) & ~15LL;
    const unsigned gb = (unsigned)std::min<int64_t>((R * kGuideK1 + 255) / 256, 65535);
    hipLaunchKernelGGL(k_encode, dim3(gb), dim3(256), 0, s, b.in, b.x, n, n_dev);
    const unsigned gw = (unsigned)std::min<int64_t>(((R / 16) + 3) / 4 + 1, 8192);
    const float *p = g->params;
    hipLaunchKernelGGL((k_gemm<kGuideK1, kGuideH, true, true>), dim3(gw), dim3(256), 0, s, b.x, g->wf[0],
                       p + layer_offset(0) + kGuideK1 * kGuideH, (void *)b.h[0], kGuideH, (const __bf16 *)nullptr, n,
                       n_dev);
    hipLaunchKernelGGL((k_gemm<kGuideH, kGuideH, true, true>), dim3(gw), dim3(256), 0, s, b.h[0], g->wf[1],
                       p + layer_offset(1) + kGuideH * kGuideH, (void *)b.h[1], kGuideH, (const __bf16 *)nullptr, n,
                       n_dev);
    hipLaunchKernelGGL((k_gemm<kGuideH, kGuideH, true, true>), dim3(gw), dim3(256), 0, s, b.h[1], g->wf[2],
                       p + layer_offset(2) + kGuideH * kGuideH, (void *)b.h[2], kGuideH, (const __bf16 *)nullptr, n,
                       n_dev);
    hipLaunchKernelGGL((k_gemm<kGuideH, kGuideNOutPad16, false, false>), dim3(gw), dim3(256), 0, s, b.h[2], g->wf[3],
                       p + layer_offset(3) + kGuideH * kGuideNOutPad16, (void *)b.raw, kGuideNOutPad16,
                       (const __bf16 *)nullptr, n, n_dev);
    return hipGetLastError();
}

// Inference over the first n rows of g->batch.in (device count n_dev optional):
// the fused kernel, raw outputs (fp32, 48 wide) into g->batch.raw.
hipError_t guide_infer(tt_guide *g, const float *in, float *raw, int64_t n, const int32_t *n_dev, hipStream_t s,
                       const int32_t *slot_of, bool store_acts) {
    InferArgs a{};
    a.in = in;
    a.slot_of = slot_of;
    if (store_acts)   // the training forward pass: hidden activations into the batch (the
                      // first layer's weight gradient recomputes the features)
        for (int k = 0; k < 3; ++k) a.h_out[k] = g->batch.h[k];
    for (int k = 0; k < 4; ++k) {
        a.w[k] = g->wf[k];
        a.bias[k] = g->params + layer_offset(k) + (int64_t)kGuideLayers[k].in * kGuideLayers[k].out;
    }
    a.raw = raw;
    a.n = n;
    a.n_dev = n_dev;
    // one block per CU (its LDS holds the weights and kInferWaves tiles); the limit is
    // set on every call (a host-side call of microseconds), so it holds on whichever
    // device and thread launches
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_mlp_infer),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, kInferLds);
    if (e != hipSuccess) return e;
    const int64_t tiles = (n + 15) / 16;
    const int64_t grid = std::max<int64_t>(
        1, std::min<int64_t>((tiles + kInferWaves - 1) / kInferWaves, std::max(1, g->ctx->num_cus)));
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)grid), dim3(64 * kInferWaves), kInferLds, s, a);
    return hipGetLastError();
}

// Gradient of the loss over rows [0, n) of the batch (inputs, dirs, wts staged).
// fused: the forward pass in the fused inference kernel (writing the activations), the
// backward products with the weight gradients of layers 4..2 in k_bwd_wgrad, and the
// first layer's weight gradient; otherwise k_encode, one k_gemm launch per product and
// one k_wgrad launch per layer (the same bits; kept as the check of the fused path).
static hipError_t guide_backward(tt_guide *g, int64_t n, hipStream_t s, bool fused = true) {
    GuideBatch &b = g->batch;
    const int64_t R = (n + 15) & ~15LL;
    hipError_t e = fused ? guide_infer(g, b.in, b.raw, n, nullptr, s, nullptr, true) : guide_forward(g, n, nullptr, s);
    if (e != hipSuccess) return e;
    // loss and weight sum: per-block partials added in a fixed order (no float atomics);
    // the weight-sum division of the gradient happens in k_reduce_all
    const unsigned lb = (unsigned)((R + kLossRows - 1) / kLossRows);
    hipLaunchKernelGGL(k_loss_grad, dim3(lb), dim3(kLossRows), 0, s, b.raw, b.dirs, b.wts,
                       g->train_alpha ? (const float *)b.pb : (const float *)nullptr, b.draw, b.lpart,
                       b.lpart + lb, n);
    hipLaunchKernelGGL(k_loss_parts, dim3(1), dim3(256), 0, s, (const float *)b.lpart, (const float *)(b.lpart + lb),
                       (int64_t)lb, g->scalars);
    // weight gradients: dW_k = H_{k-1}^T dH_k and bias gradients, partial sums per
    // block (at most kWgBlocks blocks of whole kWgRows chunks), reduced in block order
    constexpr int64_t kWgBlocks = 512, kWgMinRows = 512;
    const int64_t rpb =
        std::max<int64_t>(kWgMinRows, ((R + kWgBlocks - 1) / kWgBlocks + kWgRows - 1) / kWgRows * kWgRows);
    const int nb = (int)((R + rpb - 1) / rpb);
    struct L {
        const __bf16 *h;
        int ldh;
        const __bf16 *d;
        int ldd;
        int ki, ko;
    } ls[4] = {{b.x, kGuideK1, b.dh[0], kGuideH, kGuideK1, kGuideH},
               {b.h[0], kGuideH, b.dh[1], kGuideH, kGuideH, kGuideH},
               {b.h[1], kGuideH, b.dh[2], kGuideH, kGuideH, kGuideH},
               {b.h[2], kGuideH, b.draw, kGuideNOutPad, kGuideH, kGuideNOutPad16}};
    // one partial region per layer's weights and per layer's biases, so that a single
    // launch reduces them all
    RedRegions rg{};
    int64_t at = 0;
    int first = 0;
    for (int k = 0; k < 4; ++k) {
        for (int part = 0; part < 2; ++part) {
            const int r = 2 * k + part;
            rg.count[r] = part == 0 ? ls[k].ki * ls[k].ko : ls[k].ko;
            rg.start[r] = at;
            rg.first[r] = first;
            rg.dst[r] = (int)layer_offset(k) + (part == 0 ? 0 : ls[k].ki * ls[k].ko);
            rg.ki[r] = ls[k].ki;
            rg.ko[r] = ls[k].ko;
            rg.transpose[r] = part == 0;
            at += (int64_t)nb * rg.count[r];
            first += rg.count[r];
        }
    }
    const int64_t need = at;
    if (g->partial_cap < need) {
        if (g->partial) hipFree(g->partial);
        g->partial = nullptr;
        g->partial_cap = 0;
        if (dev_alloc(g->ctx, (void **)&g->partial, (size_t)need * 4, "guide partials") != TT_OK)
            return hipErrorOutOfMemory;
        g->partial_cap = need;
    }
    if (fused) {
        // dH3, dH2, dH1 and the weight gradients of layers 4, 3, 2 in one launch; then
        // the first layer's, its features recomputed from the inputs
        BwdArgs ba{};
        ba.draw = b.draw;
        for (int k = 0; k < 3; ++k) {
            ba.wb[k] = g->wb[k + 1];
            ba.h[k] = b.h[k];
            ba.pw[k] = g->partial + rg.start[2 * (k + 1)];
            ba.pb[k] = g->partial + rg.start[2 * (k + 1) + 1];
        }
        ba.dh1 = b.dh[0];
        ba.rows = R;
        ba.rows_per_block = rpb;
        using BS = BwShape<kBwRows, kBwWLds>;
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_bwd_wgrad<kBwRows, kBwWLds>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, BS::LDS);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_bwd_wgrad<kBwRows, kBwWLds>), dim3(nb), dim3(64 * BS::WAVES), BS::LDS, s, ba);
        const L &l = ls[0];
        hipLaunchKernelGGL((k_wgrad<kGuideK1, kGuideH, true>), dim3(nb), dim3(256), 0, s, l.h, l.ldh, l.d, l.ldd, R,
                           rpb, g->partial + rg.start[0], g->partial + rg.start[1], (const float *)b.in, n);
    } else {
        // dH3 = dRaw W4 (mask H3 > 0), dH2 = dH3 W3 (mask), dH1 = dH2 W2 (mask), one
        // launch per product, then one weight-gradient launch per layer
        const unsigned gw = (unsigned)std::min<int64_t>(((R / 16) + 3) / 4 + 1, 8192);
        hipLaunchKernelGGL((k_gemm<kGuideNOutPad, kGuideH, false, true>), dim3(gw), dim3(256), 0, s, b.draw, g->wb[3],
                           (const float *)nullptr, (void *)b.dh[2], kGuideH, (const __bf16 *)b.h[2], R,
                           (const int32_t *)nullptr);
        hipLaunchKernelGGL((k_gemm<kGuideH, kGuideH, false, true>), dim3(gw), dim3(256), 0, s, b.dh[2], g->wb[2],
                           (const float *)nullptr, (void *)b.dh[1], kGuideH, (const __bf16 *)b.h[1], R,
                           (const int32_t *)nullptr);
        hipLaunchKernelGGL((k_gemm<kGuideH, kGuideH, false, true>), dim3(gw), dim3(256), 0, s, b.dh[1], g->wb[1],
                           (const float *)nullptr, (void *)b.dh[0], kGuideH, (const __bf16 *)b.h[0], R,
                           (const int32_t *)nullptr);
        for (int k = 0; k < 4; ++k) {
            const L &l = ls[k];
            float *pw = g->partial + rg.start[2 * k], *pbias = g->partial + rg.start[2 * k + 1];
            if (k == 0)
                hipLaunchKernelGGL((k_wgrad<kGuideK1, kGuideH>), dim3(nb), dim3(256), 0, s, l.h, l.ldh, l.d, l.ldd, R,
                                   rpb, pw, pbias, (const float *)nullptr, (int64_t)0);
            else if (k < 3)
                hipLaunchKernelGGL((k_wgrad<kGuideH, kGuideH>), dim3(nb), dim3(256), 0, s, l.h, l.ldh, l.d, l.ldd, R,
                                   rpb, pw, pbias, (const float *)nullptr, (int64_t)0);
            else
                hipLaunchKernelGGL((k_wgrad<kGuideH, kGuideNOutPad16>), dim3(nb), dim3(256), 0, s, l.h, l.ldh, l.d,
                                   l.ldd, R, rpb, pw, pbias, (const float *)nullptr, (int64_t)0);
        }
    }
    hipLaunchKernelGGL(k_reduce_all, dim3((first + 63) / 64), dim3(64 * kRedParts), 0, s, (const float *)g->partial,
                       nb, first, rg, (const float *)(g->scalars + 1), g->grad);
    return hipGetLastError();
}

static hipError_t guide_adam(tt_guide *g, hipStream_t s) {
    g->step++;
    const float c1 = 1.0f - std::pow(g->beta1, (float)g->step), c2 = 1.0f - std::pow(g->beta2, (float)g->step);
    const int n = (int)guide_param_count();
    AdamLayers L{};
    for (int k = 0; k < 4; ++k) {
        L.off[k] = (int)layer_offset(k);
        L.in[k] = kGuideLayers[k].in;
        L.out[k] = kGuideLayers[k].out;
        L.out_pad_bwd[k] = k == 3 ? kGuideNOutPad : kGuideLayers[k].out;
        L.fwd[k] = g->wf[k];
        L.bwd[k] = g->wb[k];
    }
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, g->params, g->m1, g->m2, g->grad, n, g->lr,
                       g->beta1, g->beta2, g->eps, c1, c2, L);
    return hipGetLastError();
}

hipError_t guide_train_rows(tt_guide *g, int64_t n, hipStream_t s) {
    hipError_t e = guide_backward(g, n, s);
    if (e == hipSuccess) e = guide_adam(g, s);
    return e;
}

}  // namespace tt

using namespace tt;

#define TT_HIPG(call, what)                                                                       \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return fail(TT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" {

tt_status tt_guide_create(tt_ctx *ctx, uint32_t seed, float lr, tt_guide **out) {
    if (!ctx || !out || !(lr > 0.0f)) return fail(TT_ERR_INVALID, "tt_guide_create: bad arguments");
    TT_HIPG(hipSetDevice(ctx->device), "hipSetDevice");
    auto *g = new (std::nothrow) tt_guide();
    if (!g) return fail(TT_ERR_NOMEM, "tt_guide_create: host allocation failed");
    g->ctx = ctx;
    g->lr = lr;
    const int64_t np = guide_param_count();
    tt_status st = TT_OK;
    auto al = [&](void **p, size_t bytes, const char *w) {
        if (st == TT_OK) st = dev_alloc(ctx, p, bytes, w);
    };
    al((void **)&g->params, (size_t)np * 4, "guide params");
    al((void **)&g->m1, (size_t)np * 4, "guide adam m");
    al((void **)&g->m2, (size_t)np * 4, "guide adam v");
    al((void **)&g->grad, (size_t)np * 4, "guide grad");
    al((void **)&g->scalars, 2 * 4, "guide scalars");
    for (int k = 0; k < 4; ++k) {
        const GuideLayer &l = kGuideLayers[k];
        al((void **)&g->wf[k], (size_t)l.in * l.out * 2, "guide bf16 weights");
        const int out_pad_bwd = k == 3 ? kGuideNOutPad : l.out;
        al((void **)&g->wb[k], (size_t)l.in * out_pad_bwd * 2, "guide bf16 weights (transposed)");
    }
    if (st != TT_OK) {
        tt_guide_destroy(g);
        return st;
    }
    // He-normal initialisation (Box-Muller over Philox), zero biases; the output
    // layer is scaled down so the initial mixture is broad and evenly weighted.
    std::vector<float> p((size_t)np, 0.0f);
    uint32_t ctr = 0;
    for (int k = 0; k < 4; ++k) {
        const GuideLayer &l = kGuideLayers[k];
        const double sd = (k == 3 ? 0.1 : 1.0) * std::sqrt(2.0 / l.in);
        float *w = p.data() + layer_offset(k);
        for (int e = 0; e < l.in * l.out; ++e) {
            u32x4 r = philox4x32_10({ctr++, 0x6A1D, 0, 0}, seed, 0x7u);
            double u1 = 1.0 - u01d(r.x, r.y), u2 = u01d(r.z, r.w);
            double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            const int o = e / l.in;
            w[e] = (k == 3 && o >= kGuideNOut) ? 0.0f : (float)(sd * z);
        }
    }
    // the learned selection probability starts low (alpha = learned_alpha(-3) = 0.064):
    // an untrained mixture is sampled rarely until training raises it where it helps
    p[(size_t)layer_offset(3) + (size_t)kGuideLayers[3].in * kGuideLayers[3].out + kGuideAlphaOut] = -3.0f;
    hipError_t e = hipMemcpy(g->params, p.data(), (size_t)np * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(g->m1, 0, (size_t)np * 4);
    if (e == hipSuccess) e = hipMemset(g->m2, 0, (size_t)np * 4);
    for (int k = 0; k < 4 && e == hipSuccess; ++k) {
        const GuideLayer &l = kGuideLayers[k];
        const int out_pad_bwd = k == 3 ? kGuideNOutPad : l.out;
        e = hipMemset(g->wb[k], 0, (size_t)l.in * out_pad_bwd * 2);
    }
    if (e == hipSuccess) e = refresh_bf16(g, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        tt_guide_destroy(g);
        return fail(TT_ERR_DEVICE, std::string("tt_guide_create: ") + hipGetErrorString(e));
    }
    *out = g;
    return TT_OK;
}

void tt_guide_destroy(tt_guide *g) {
    if (!g) return;
    hipSetDevice(g->ctx->device);
    hipStreamSynchronize(g->ctx->stream);
    GuideBatch &b = g->batch;
    void *ptrs[] = {g->params, g->m1, g->m2, g->grad, g->scalars, g->partial, b.in, b.x, b.h[0], b.h[1], b.h[2],
                    b.raw, b.draw, b.dh[0], b.dh[1], b.dh[2], b.dirs, b.wts, b.pb, b.lpart, g->wf[0], g->wf[1], g->wf[2], g->wf[3],
                    g->wb[0], g->wb[1], g->wb[2], g->wb[3], g->train_buf, g->train_off, g->rec, g->iin, g->iraw,
                    g->imask, g->iblock_n, g->islot};
    for (void *q : ptrs)
        if (q) hipFree(q);
    delete g;
}

int64_t tt_guide_param_count(void) { return guide_param_count(); }

tt_status tt_guide_get_params(tt_guide *g, float *out) {
    if (!g || !out) return fail(TT_ERR_INVALID, "tt_guide_get_params: bad arguments");
    TT_HIPG(hipMemcpy(out, g->params, (size_t)guide_param_count() * 4, hipMemcpyDeviceToHost), "copy params");
    return TT_OK;
}

tt_status tt_guide_set_params(tt_guide *g, const float *in) {
    if (!g || !in) return fail(TT_ERR_INVALID, "tt_guide_set_params: bad arguments");
    TT_HIPG(hipSetDevice(g->ctx->device), "hipSetDevice");
    TT_HIPG(hipMemcpy(g->params, in, (size_t)guide_param_count() * 4, hipMemcpyHostToDevice), "copy params");
    TT_HIPG(refresh_bf16(g, g->ctx->stream), "bf16 refresh");
    TT_HIPG(hipStreamSynchronize(g->ctx->stream), "sync");
    return TT_OK;
}

// The n x 40 outputs from the padded n x 48 rows: one contiguous copy, then the
// rows are packed on the host (a strided 2-D copy into pageable memory moves one
// row at a time and took seconds for millions of rows).
static tt_status copy_raw_out(tt_guide *g, int64_t n, float *raw_out, hipStream_t s) {
    std::vector<float> pad((size_t)n * kGuideNOutPad16);
    TT_HIPG(hipMemcpyAsync(pad.data(), g->batch.raw, pad.size() * 4, hipMemcpyDeviceToHost, s), "copy out");
    TT_HIPG(hipStreamSynchronize(s), "sync");
    for (int64_t r = 0; r < n; ++r)
        std::memcpy(raw_out + r * kGuideNOut, pad.data() + r * kGuideNOutPad16, kGuideNOut * 4);
    return TT_OK;
}

tt_status tt_guide_forward(tt_guide *g, const float *inputs, int64_t n, float *raw_out) {
    if (!g || !inputs || !raw_out || n <= 0 || n > (1LL << 24)) return fail(TT_ERR_INVALID, "tt_guide_forward: bad arguments");
    TT_HIPG(hipSetDevice(g->ctx->device), "hipSetDevice");
    tt_status st = guide_alloc_batch(g, n);
    if (st != TT_OK) return st;
    hipStream_t s = g->ctx->stream;
    TT_HIPG(hipMemcpyAsync(g->batch.in, inputs, (size_t)n * kGuideIn * 4, hipMemcpyHostToDevice, s), "copy inputs");
    TT_HIPG(guide_infer(g, g->batch.in, g->batch.raw, n, nullptr, s), "guide inference");
    return copy_raw_out(g, n, raw_out, s);
}

tt_status tt_debug_guide_forward_unfused(tt_guide *g, const float *inputs, int64_t n, float *raw_out) {
    if (!g || !inputs || !raw_out || n <= 0 || n > (1LL << 24))
        return fail(TT_ERR_INVALID, "tt_debug_guide_forward_unfused: bad arguments");
    TT_HIPG(hipSetDevice(g->ctx->device), "hipSetDevice");
    tt_status st = guide_alloc_batch(g, n);
    if (st != TT_OK) return st;
    hipStream_t s = g->ctx->stream;
    TT_HIPG(hipMemcpyAsync(g->batch.in, inputs, (size_t)n * kGuideIn * 4, hipMemcpyHostToDevice, s), "copy inputs");
    TT_HIPG(guide_forward(g, n, nullptr, s), "guide forward");
    return copy_raw_out(g, n, raw_out, s);
}

static tt_status stage_rows(tt_guide *g, const float *inputs, const float *dirs, const float *wts, int64_t n) {
    if (!g || !inputs || !dirs || !wts || n <= 0 || n > (1LL << 24)) return fail(TT_ERR_INVALID, "guide: bad arguments");
    if (hipSetDevice(g->ctx->device) != hipSuccess) return fail(TT_ERR_DEVICE, "hipSetDevice failed");
    tt_status st = guide_alloc_batch(g, n);
    if (st != TT_OK) return st;
    hipStream_t s = g->ctx->stream;
    hipError_t e = hipMemcpyAsync(g->batch.in, inputs, (size_t)n * kGuideIn * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(g->batch.dirs, dirs, (size_t)n * 12, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(g->batch.wts, wts, (size_t)n * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return fail(TT_ERR_DEVICE, std::string("guide staging: ") + hipGetErrorString(e));
    return TT_OK;
}

static tt_status grad_impl(tt_guide *g, const float *inputs, const float *dirs, const float *wts, int64_t n,
                           float *loss, float *grad_out, bool fused) {
    tt_status st = stage_rows(g, inputs, dirs, wts, n);
    if (st != TT_OK) return st;
    hipStream_t s = g->ctx->stream;
    TT_HIPG(guide_backward(g, n, s, fused), "guide backward");
    float sc[2];
    TT_HIPG(hipMemcpyAsync(sc, g->scalars, 8, hipMemcpyDeviceToHost, s), "copy loss");
    if (grad_out)
        TT_HIPG(hipMemcpyAsync(grad_out, g->grad, (size_t)guide_param_count() * 4, hipMemcpyDeviceToHost, s), "copy grad");
    TT_HIPG(hipStreamSynchronize(s), "sync");
    if (loss) *loss = sc[0];
    return TT_OK;
}

tt_status tt_guide_grad(tt_guide *g, const float *inputs, const float *dirs, const float *wts, int64_t n, float *loss,
                        float *grad_out) {
    return grad_impl(g, inputs, dirs, wts, n, loss, grad_out, true);
}

tt_status tt_debug_guide_grad_unfused(tt_guide *g, const float *inputs, const float *dirs, const float *wts, int64_t n,
                                      float *loss, float *grad_out) {
    return grad_impl(g, inputs, dirs, wts, n, loss, grad_out, false);
}

tt_status tt_guide_train_step(tt_guide *g, const float *inputs, const float *dirs, const float *wts, int64_t n,
                              float *loss) {
    tt_status st = stage_rows(g, inputs, dirs, wts, n);
    if (st != TT_OK) return st;
    hipStream_t s = g->ctx->stream;
    TT_HIPG(guide_train_rows(g, n, s), "guide train step");
    float sc[2];
    TT_HIPG(hipMemcpyAsync(sc, g->scalars, 8, hipMemcpyDeviceToHost, s), "copy loss");
    TT_HIPG(hipStreamSynchronize(s), "sync");
    if (loss) *loss = sc[0];
    return TT_OK;
}

}  // extern "C"
