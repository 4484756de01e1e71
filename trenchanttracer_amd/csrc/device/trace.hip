// Closest-hit / any-hit traversal of the SBVH on gfx950 (wave64).
//
// Per lane: one ray at a time (a lane whose ray is done takes the next one of its
// wave's sequence), the while-while loop of Aila & Laine ("Understanding the
// Efficiency of Ray Traversal on GPUs", HPG 2009) in the form of the reference
// intersect() (include/geometry/IntersectKernel.cuh:78-296):
//   - ray precompute idir / ood with the 2^-80 clamp                (:92-101)
//   - two child boxes per 64-byte node, slab entry/exit as 3-operand integer
//     min/max of the float bit patterns (v_min3_i32 / v_max3_i32 restate the
//     PTX vmin/vmax.s32 forms), traverse iff entry <= exit, entry >= tmin,
//     entry <= 1e20                                                  (:126-147)
//   - near child first, far child pushed                             (:149-167)
//   - first leaf postponed                                           (:172-178)
//   - Möller–Trumbore on v0 and the two edges, front faces only with
//     |det| > 1e-5, tmin < t < best                                  (:230-257)
// Arithmetic is unfused (fp contract off) and division IEEE, so for a given
// visiting order every decision is bit-identical to the host restatement.
//
// Two loop variants:
//   VARIANT 0 (exact): a lane leaves the node loop as soon as it holds a leaf,
//     i.e. the single-lane order. Hit ids, t, u, v are bit-identical to the
//     host traversal of the same tree.
//   VARIANT 1 (speculative): lanes keep descending until no lane of the wave is
//     still searching for its first leaf (__any ballot, :192-203). Results differ
//     from the single-lane order only where a triangle lies on a box boundary
//     within rounding of the current hit distance.
// Stack: the top kLds entries of each lane live in LDS ([entry][lane] so the
// lanes of a wave hit distinct banks), deeper entries in per-lane scratch.
#include "dev_common.h"

#pragma clang fp contract(off)

namespace tt {
namespace {

__device__ __forceinline__ int fb(float f) { return __float_as_int(f); }
__device__ __forceinline__ float bf(int i) { return __int_as_float(i); }

// fminf / fmaxf of the slab code with -0 < +0 (v_minimum3_f32 / v_maximum3_f32).
__device__ __forceinline__ float fmn(float a, float b) { return __builtin_elementwise_minimum(a, b); }
__device__ __forceinline__ float fmx(float a, float b) { return __builtin_elementwise_maximum(a, b); }

// spanBeginKepler / spanEndKepler (IntersectKernel.cuh:54-60).
__device__ __forceinline__ float span_begin(float a0, float a1, float b0, float b1, float c0, float c1, float d) {
    int z = max(min(fb(c0), fb(c1)), fb(d));
    return bf(max(max(fb(fmn(a0, a1)), fb(fmn(b0, b1))), z));
}
__device__ __forceinline__ float span_end(float a0, float a1, float b0, float b1, float c0, float c1, float d) {
    int z = min(max(fb(c0), fb(c1)), fb(d));
    return bf(min(min(fb(fmx(a0, a1)), fb(fmx(b0, b1))), z));
}

// The overflow array lives outside the struct: a struct that holds a dynamically
// indexed array is kept in scratch as a whole, and then every push and pop would
// load and store sp and the LDS pointer through scratch as well.
template <int S>
struct Stack {
    int *lds;       // this lane's column: lds[k * stride]
    int stride;
    int sp;         // entries held (the sentinel lies below entry 0)
    int off;        // sp * stride, kept by additions (no multiply per push / pop)
    int *spill;     // this lane's private overflow, kStackCap - S entries

    __device__ __forceinline__ void push(int v) {
        if (sp < S) {
            lds[off] = v;
        } else if (sp < kStackCap) {
            spill[sp - S] = v;
        }
        sp++;
        off += stride;
    }
    __device__ __forceinline__ int pop() {
        if (sp <= 0) return kSentinel;
        sp--;
        off -= stride;
        if (sp < S) return lds[off];
        if (sp < kStackCap) return spill[sp - S];
        return kSentinel;
    }
};

// REC (closest hit only; the node-fetch probe): 1 counts each ray's node visits,
// triangle tests and leaf visits; 2 records at every node visit the slab test's
// outcome as a 4-bit code (bit 0: child 0 entered, bit 1: child 1 entered, bit 2:
// both, child 1 first); 3 replays: the same loop, stack, leaf postponement, ballots
// and refills, every node and triangle record fetched, but each node's outcome read
// from the record instead of computed and no triangle arithmetic.
template <int VARIANT, bool ANY, int S, int TRI, int REC = 0>
__global__ __launch_bounds__(256) void k_trace(TraceArgs a) {
    static_assert(REC == 0 || !ANY, "fetch recording follows the closest-hit loop");
    extern __shared__ int lds_stack[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // XCD-contiguous order: the hardware hands blocks to the 8 XCDs round-robin, so
    // block b runs on XCD b % 8; renumbering b -> (b % 8) * q + b / 8 (for the first
    // 8q blocks, q = blocks / 8; the tail keeps its number) is a bijection that gives
    // each XCD's L2 one contiguous range of waves
    unsigned vb = blockIdx.x;
    if (a.xcd_order) {
        const unsigned q = gridDim.x >> 3;
        if (vb < (q << 3)) vb = (vb & 7u) * q + (vb >> 3);
    }
    int64_t n = a.n;
    if (a.n_dev) n = min((int64_t)*a.n_dev, n);
    if (blockIdx.x == 0 && tid == 0) {
        if (a.reset0) *a.reset0 = 0;
        if (a.reset1) *a.reset1 = 0;
    }
    // Persistent waves (Aila & Laine's persistent threads with ray replacement):
    // the grid holds about as many waves as the chip keeps resident, and wave w of W
    // owns the 64-ray groups w, w + W, w + 2W, ... in that order, so that all waves
    // sweep the ray range together (the rays in flight at any moment lie in a window
    // of about 2W groups, as with one block per group, which keeps coherent rays'
    // nodes in cache). A lane whose ray is done writes its result and takes the next
    // ray of its wave's sequence (ballot and prefix count, no atomics): the lanes
    // stay busy until the sequence runs out, instead of idling behind the wave's
    // longest ray. Each ray's traversal is the same as with one ray per lane.
    const int64_t groups = (n + 63) >> 6;
    const int64_t waves_all = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t waves = max((int64_t)1, min(waves_all, groups));
    const int64_t w = (int64_t)vb * (blockDim.x >> 6) + (tid >> 6);
    if (w >= waves) return;   // whole waves only: nothing below synchronises the block
    const int64_t total = ((groups - w + waves - 1) / waves) * 64;   // the wave's sequence positions
    const unsigned long long below = (1ull << lane) - 1ull;

    int spill[kStackCap - S];
    Stack<S> st;
    st.lds = lds_stack + tid;
    st.stride = blockDim.x;
    st.sp = 0;
    st.off = 0;
    st.spill = spill;

    int64_t rid = -1;          // the lane's ray (-1: none)
    int64_t cursor = 0;        // next sequence position (uniform)
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, tmin = 0.0f, dx = 0.0f, dy = 0.0f, dz = 1.0f;
    float idirx = 0.0f, idiry = 0.0f, idirz = 0.0f, oodx = 0.0f, oody = 0.0f, oodz = 0.0f;
    int leaf_addr = 0;
    int node_addr = kSentinel;
    float hit_t = 0.0f, hit_u = 0.0f, hit_v = 0.0f;
    int hit_id = -1;
    int rec_n = 0, rec_t = 0, rec_l = 0;   // node visits, triangle tests, leaf visits of the lane's ray (REC)
    int64_t rec_base = 0, rec_cap = 0;     // the ray's words of outcome codes (REC 2, 3)
    uint32_t rec_word = 0;                 // codes of the current 8 visits
    uint32_t rec_next = 0;                 // codes of the next 8 (REC 3: loaded a word ahead)
    uint32_t tok = 0;                      // every fetched word folded in (REC 3)

    for (;;) {
        const bool done = node_addr == kSentinel;
        if (done && rid >= 0) {
            if (REC == 1) {
                a.rec_nodes[rid] = rec_n;
                a.rec_tris[rid] = rec_t;
                a.rec_leaves[rid] = rec_l;
            } else if (REC == 2) {
                if ((rec_n & 7) != 0 && (rec_n >> 3) < rec_cap) a.rec_log[rec_base + (rec_n >> 3)] = (int32_t)rec_word;
            } else if (REC == 3) {
                *reinterpret_cast<float4 *>(a.hits + rid) = make_float4(bf((int)tok), 0.0f, 0.0f, 0.0f);
            } else if (ANY) {
                a.flags[rid] = hit_id != -1 ? 1 : 0;
            } else {
                if (hit_id == -1) {
                    hit_u = 0.0f;
                    hit_v = 0.0f;
                }
                *reinterpret_cast<float4 *>(a.hits + rid) = make_float4(bf(hit_id), hit_t, hit_u, hit_v);
            }
            rid = -1;
        }
        const unsigned long long m = __ballot(done);
        // new rays once enough lanes are idle (all idle always qualifies, so the
        // wave cannot wait on itself)
        const bool take = (int32_t)__popcll(m) >= a.refill;
        if (take && done) {
            const int64_t q = cursor + __popcll(m & below);
            if (q < total) {
                const int64_t r = ((q >> 6) * waves + w) * 64 + (q & 63);
                if (r < n) rid = a.order ? (int64_t)a.order[r] : r;
            }
            if (rid >= 0) {
                const float4 *rp = reinterpret_cast<const float4 *>(a.rays + rid);
                const float4 r0 = rp[0], r1 = rp[1];
                ox = r0.x;
                oy = r0.y;
                oz = r0.z;
                tmin = r0.w;
                dx = r1.x;
                dy = r1.y;
                dz = r1.z;
                const float ooeps = 0x1p-80f;
                idirx = 1.0f / (fabsf(dx) > ooeps ? dx : copysignf(ooeps, dx));
                idiry = 1.0f / (fabsf(dy) > ooeps ? dy : copysignf(ooeps, dy));
                idirz = 1.0f / (fabsf(dz) > ooeps ? dz : copysignf(ooeps, dz));
                oodx = ox * idirx;
                oody = oy * idiry;
                oodz = oz * idirz;
                st.sp = 0;
                st.off = 0;
                leaf_addr = 0;
                node_addr = a.num_nodes > 0 ? (a.ray_root ? a.ray_root[rid] : 0) : kSentinel;
                hit_t = r1.w;
                hit_u = 0.0f;
                hit_v = 0.0f;
                hit_id = -1;
                rec_n = 0;
                rec_t = 0;
                rec_l = 0;
                rec_word = 0;
                if (REC >= 2) {
                    rec_base = a.rec_off[rid];
                    rec_cap = a.rec_off[rid + 1] - rec_base;
                }
                if (REC == 3) {
                    tok = __float_as_uint(r0.x) ^ __float_as_uint(r1.w);
                    rec_next = rec_cap > 0 ? (uint32_t)a.rec_log[rec_base] : 0u;
                }
            }
        }
        if (take) cursor += __popcll(m);
        // ends once the sequence is used up and no lane holds a ray (only the last
        // group of the range can be partial)
        if (cursor >= total && !__any(rid >= 0)) break;

        while (node_addr >= 0 && node_addr != kSentinel) {
            const float4 *np = a.nodes + (int64_t)node_addr * 4;
            const float4 n0xy = np[0];
            const float4 n1xy = np[1];
            const float4 nz = np[2];
            const int2 ch = reinterpret_cast<const int2 *>(np + 3)[0];
            bool t0, t1, c1_first;
            if (REC == 3) {
                // the recorded outcome; the box words are still fetched and used
                if ((rec_n & 7) == 0) {
                    // the next word is requested 8 visits before it is needed
                    rec_word = rec_next;
                    const int64_t wn = (rec_n >> 3) + 1;
                    rec_next = wn < rec_cap ? (uint32_t)a.rec_log[rec_base + wn] : 0u;
                }
                const uint32_t code = (rec_word >> ((rec_n & 7) * 4)) & 15u;
                ++rec_n;
                t0 = (code & 1u) != 0;
                t1 = (code & 2u) != 0;
                c1_first = (code & 4u) != 0;
                tok ^= (__float_as_uint(n0xy.x) ^ __float_as_uint(n0xy.y) ^ __float_as_uint(n0xy.z) ^
                        __float_as_uint(n0xy.w)) ^
                       (__float_as_uint(n1xy.x) ^ __float_as_uint(n1xy.y) ^ __float_as_uint(n1xy.z) ^
                        __float_as_uint(n1xy.w)) ^
                       (__float_as_uint(nz.x) ^ __float_as_uint(nz.y) ^ __float_as_uint(nz.z) ^ __float_as_uint(nz.w));
            } else {

            const float c0lox = n0xy.x * idirx - oodx;
            const float c0hix = n0xy.y * idirx - oodx;
            const float c0loy = n0xy.z * idiry - oody;
            const float c0hiy = n0xy.w * idiry - oody;
            const float c0loz = nz.x * idirz - oodz;
            const float c0hiz = nz.y * idirz - oodz;
            const float c1loz = nz.z * idirz - oodz;
            const float c1hiz = nz.w * idirz - oodz;
            const float c0min = span_begin(c0lox, c0hix, c0loy, c0hiy, c0loz, c0hiz, tmin);
            const float c0max = span_end(c0lox, c0hix, c0loy, c0hiy, c0loz, c0hiz, hit_t);
            const float c1lox = n1xy.x * idirx - oodx;
            const float c1hix = n1xy.y * idirx - oodx;
            const float c1loy = n1xy.z * idiry - oody;
            const float c1hiy = n1xy.w * idiry - oody;
            const float c1min = span_begin(c1lox, c1hix, c1loy, c1hiy, c1loz, c1hiz, tmin);
            const float c1max = span_end(c1lox, c1hix, c1loy, c1hiy, c1loz, c1hiz, hit_t);

            const float ray_tmax = 1e20f;
            t0 = (c0min <= c0max) && (c0min >= tmin) && (c0min <= ray_tmax);
            t1 = (c1min <= c1max) && (c1min >= tmin) && (c1min <= ray_tmax);
            c1_first = c1min < c0min;
            if (REC == 2) {
                const uint32_t code = (t0 ? 1u : 0u) | (t1 ? 2u : 0u) | ((t0 && t1 && c1_first) ? 4u : 0u);
                rec_word |= code << ((rec_n & 7) * 4);
                if ((rec_n & 7) == 7) {
                    if ((rec_n >> 3) < rec_cap) a.rec_log[rec_base + (rec_n >> 3)] = (int32_t)rec_word;
                    rec_word = 0;
                }
            }
            if (REC) ++rec_n;
            }
            // the child indices are used only when a child is entered; without a use on
            // every path the compiler issues their load there, after the slab tests,
            // which makes it a second dependent fetch per visited node. This empty
            // statement uses them here, so all 64 bytes of the node are fetched at once.
            asm volatile("" : : "v"(ch.x), "v"(ch.y));

            if (!t0 && !t1) {
                node_addr = st.pop();
            } else {
                int far = ch.y;
                node_addr = t0 ? ch.x : ch.y;
                if (t0 && t1) {
                    if (c1_first) {
                        int tmp = node_addr;
                        node_addr = far;
                        far = tmp;
                    }
                    st.push(far);
                }
            }
            if (node_addr < 0 && leaf_addr >= 0) {
                leaf_addr = node_addr;
                node_addr = st.pop();
            }
            if (VARIANT == 0) {
                if (leaf_addr < 0) break;
            } else {
                if (!__any(leaf_addr >= 0)) break;
            }
        }

        while (leaf_addr < 0) {
            int tri = ~leaf_addr;
            if (REC) ++rec_l;
            for (;;) {
                if (REC) ++rec_t;
                const float4 *tp = a.tris + (int64_t)tri * 3;
                const float4 A = tp[0];
                const float4 B = tp[1];
                const float4 C = tp[2];
                if (REC == 3) {
                    // the record is fetched and its words used; no test
                    tok ^= (__float_as_uint(A.x) ^ __float_as_uint(A.y) ^ __float_as_uint(A.z) ^ __float_as_uint(A.w)) ^
                           (__float_as_uint(B.x) ^ __float_as_uint(B.y) ^ __float_as_uint(B.z) ^ __float_as_uint(B.w)) ^
                           (__float_as_uint(C.x) ^ __float_as_uint(C.y) ^ __float_as_uint(C.z));
                    if (fb(C.w) != 0) break;
                    ++tri;
                    continue;
                }
                // Möller–Trumbore, IntersectKernel.cuh:230-254 (edges precomputed: same bits)
                const float tvx = ox - A.x, tvy = oy - A.y, tvz = oz - A.z;
                const float pvx = dy * C.z - dz * C.y;
                const float pvy = dz * C.x - dx * C.z;
                const float pvz = dx * C.y - dy * C.x;
                const float det = B.x * pvx + B.y * pvy + B.z * pvz;
                const float inv = 1.0f / det;
                const float u = (tvx * pvx + tvy * pvy + tvz * pvz) * inv;
                const float qvx = tvy * B.z - tvz * B.y;
                const float qvy = tvz * B.x - tvx * B.z;
                const float qvz = tvx * B.y - tvy * B.x;
                const float v = (dx * qvx + dy * qvy + dz * qvz) * inv;
                bool ok;
                if (TRI == kTriReference) {
                    const float EPS = 0.00001f;
                    if (det > EPS) {
                        ok = !(u < 0.0f || u > 1.0f) && !(v < 0.0f || (u + v) > 1.0f);
                    } else if (det < -EPS) {
                        // the reference's back-face bounds (IntersectKernel.cuh:248-250)
                        // have their comparisons reversed: u > 0 || u < 1 always holds,
                        // so back faces are never accepted; kept as written
                        ok = !(u > 0.0f || u < 1.0f) && !(v > 0.0f || (u + v) < 1.0f);
                    } else {
                        ok = false;
                    }
                } else {
                    ok = det != 0.0f && u >= 0.0f && u <= 1.0f && v >= 0.0f && (u + v) <= 1.0f;
                }
                if (ok) {
                    const float t = (C.x * qvx + C.y * qvy + C.z * qvz) * inv;
                    if (tmin < t && t < hit_t) {
                        hit_t = t;
                        hit_u = u;
                        hit_v = v;
                        hit_id = fb(A.w);
                        if (ANY) {
                            node_addr = kSentinel;
                            break;
                        }
                    }
                }
                if (fb(C.w) != 0) break;
                ++tri;
            }
            leaf_addr = node_addr;
            if (node_addr < 0) node_addr = st.pop();
        }
    }
}

// Persistent launch: as many blocks as the CUs hold at once (occupancy of this
// instantiation at this block size and LDS), and for a known ray count no more
// blocks than 64-ray groups. Otherwise one wave per 64-ray group (every wave's
// sequence is its one group): the hardware dispatcher then balances the waves,
// which wins where rays are coherent or the count is known.
template <class K>
hipError_t launch_k(K kernel, const TraceArgs &a, bool persistent, int block, size_t lds, int num_cus, hipStream_t s) {
    int64_t grid = (a.n + block - 1) / block;
    if (persistent) {
        int per_cu = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds);
        if (e != hipSuccess) return e;
        int64_t resident = (int64_t)std::max(1, per_cu) * std::max(1, num_cus);
        grid = a.n_dev ? resident : std::min<int64_t>(resident, (a.n + 63) / 64);
    }
    grid = std::max<int64_t>(1, grid);
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(block), lds, s, a);
    return hipGetLastError();
}

template <int V, bool ANY, int TRI>
hipError_t launch_s(const TraceArgs &a, bool persistent, int S, int block, int num_cus, hipStream_t s) {
    const size_t lds = (size_t)S * block * sizeof(int);
    switch (S) {
        case 8: return launch_k(k_trace<V, ANY, 8, TRI>, a, persistent, block, lds, num_cus, s);
        case 16: return launch_k(k_trace<V, ANY, 16, TRI>, a, persistent, block, lds, num_cus, s);
        case 24: return launch_k(k_trace<V, ANY, 24, TRI>, a, persistent, block, lds, num_cus, s);
        case 32: return launch_k(k_trace<V, ANY, 32, TRI>, a, persistent, block, lds, num_cus, s);
        default: return hipErrorInvalidValue;
    }
}

template <int TRI>
hipError_t launch_v(const TraceArgs &a, bool persistent, int variant, bool any, int S, int block, int num_cus,
                    hipStream_t s) {
    if (variant == 0)
        return any ? launch_s<0, true, TRI>(a, persistent, S, block, num_cus, s)
                   : launch_s<0, false, TRI>(a, persistent, S, block, num_cus, s);
    if (variant == 1)
        return any ? launch_s<1, true, TRI>(a, persistent, S, block, num_cus, s)
                   : launch_s<1, false, TRI>(a, persistent, S, block, num_cus, s);
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_trace_record(const TraceArgs &a, int variant, int pass, int block, int refill, int num_cus,
                               hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    if (block < 64 || block > 256 || block % 64 != 0) return hipErrorInvalidValue;
    TraceArgs b = a;
    b.xcd_order = 0;
    const bool persistent = refill > 0;
    b.refill = persistent ? (refill > 64 ? 64 : refill) : 64;
    const size_t lds = (size_t)16 * block * sizeof(int);
#define TT_REC(V, P) \
    if (variant == V && pass == P) return launch_k(k_trace<V, false, 16, kTriReference, P>, b, persistent, block, lds, num_cus, s);
    TT_REC(0, 1) TT_REC(0, 2) TT_REC(0, 3) TT_REC(1, 1) TT_REC(1, 2) TT_REC(1, 3)
#undef TT_REC
    return hipErrorInvalidValue;
}

hipError_t launch_trace(const TraceArgs &a, int variant, bool any, int lds_stack, int block, int xcd_order,
                        int refill, int num_cus, hipStream_t s, int tri_mode) {
    if (a.n <= 0) return hipSuccess;
    if (block < 64 || block > 256 || block % 64 != 0) return hipErrorInvalidValue;
    if ((a.n + block - 1) / block > 0x7FFFFFFF) return hipErrorInvalidValue;
    TraceArgs b = a;
    b.xcd_order = xcd_order ? 1 : 0;
    // refill 0: one wave per 64-ray group; 1..64: persistent waves with that threshold
    const bool persistent = refill > 0;
    b.refill = persistent ? (refill > 64 ? 64 : refill) : 64;
    if (tri_mode == kTriReference) return launch_v<kTriReference>(b, persistent, variant, any, lds_stack, block, num_cus, s);
    if (tri_mode == kTriTwoSided) return launch_v<kTriTwoSided>(b, persistent, variant, any, lds_stack, block, num_cus, s);
    return hipErrorInvalidValue;
}

}  // namespace tt
