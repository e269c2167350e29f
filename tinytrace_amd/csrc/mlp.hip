// Radiance-cache MLP on gfx950 matrix cores.
//
// Every matmul is D[M x N] = A[M x K] B[K x N] with
// v_mfma_f32_32x32x8bf16_1k: lane l holds A row (l % 32), B column (l % 32),
// k = 4 (l / 32) + 0..3 of each 8-deep step, and D column (l % 32), rows
// 8 (r / 4) + 4 (l / 32) + (r % 4) for its 16 accumulators r.  Samples are
// the N dimension (32 per wave tile) and features the M / K dimensions:
// Z = W A_prev.  Then the four accumulators r = 4g .. 4g+3 of output tile t
// are exactly the B operand of k-step 4t + g of the next layer, so a sample
// tile runs every layer with its activations in registers.  Weights sit in
// LDS as bf16 rows padded by 4 halves (conflict-free for the 32 rows a wave
// reads).
//
// Training, per Adam step (two launches; the record count is read on the
// device, so a render enqueues its training without a host round trip):
//   mlp_train_kernel  one workgroup per 256-sample chunk, one 32-sample tile
//                     per wave: forward + backward in registers, activations
//                     and deltas staged (bf16) in the chunk's own plane, then
//                     dW_l = delta_l a_l^T over the chunk into the chunk's
//                     fp32 partial, and the chunk's loss sum
//   mlp_adam_kernel   sums the partials in chunk order (deterministic),
//                     applies Adam to the fp32 master weights and writes the
//                     updated weights' bf16 copies
// Training runs on the context's API stream, after an event of the render
// whose records it reads; its `trained` event orders later inference.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "mlp.h"
#include "abi_guard.h"
#include "runtime_internal.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

using namespace tt;

namespace {

constexpr int MLP_BLOCK = 256;                 // inference: 4 waves
constexpr int TRAIN_WAVES = 8;
constexpr int TRAIN_BLOCK = 64 * TRAIN_WAVES;
constexpr int MLP_CHUNK = 32 * TRAIN_WAVES;    // samples per training workgroup
constexpr int IN_STEPS = MLP_IN / 8;           // k-steps of layer 0
// A chunk's staging plane, rows of MLP_CHUNK samples (bf16): a_0 (96 rows),
// a_1 .. a_5 (64 each), then d_0 .. d_4 (64 each) and d_5 (3), where
// d_k = dLoss / d(W_k a_k).
constexpr int ACT_ROWS = MLP_IN + 5 * 64;      // 416
constexpr int PLANE_ROWS = ACT_ROWS + 5 * 64 + 3;
constexpr uint32_t MAX_STEPS = 64;             // Adam steps per training call
constexpr int WGRAD_TILES = 6 + 16 + 2;        // 32 x 32 tiles of dW_0, dW_1..4, dW_5

__host__ __device__ constexpr int act_row(int k) { return k == 0 ? 0 : MLP_IN + 64 * (k - 1); }
__host__ __device__ constexpr int delta_row(int k) { return ACT_ROWS + 64 * k; }

__device__ __forceinline__ f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// fp32 -> bf16 with the hardware convert (v_cvt_pk_bf16_f32, round to
// nearest even, as float_to_bf16).
__device__ __forceinline__ short bf(float x) { return __builtin_bit_cast(short, static_cast<__bf16>(x)); }
__device__ __forceinline__ float unbf(short x) { return bf16_to_float(uint16_t(x)); }

// A operand: rows row0 + (l % 32) of a row-major bf16 matrix with `stride`
// halves per row, k-step s.
__device__ __forceinline__ bf16x4 load_a(const uint16_t* m, int stride, int row0, int s, uint32_t l) {
    return *reinterpret_cast<const bf16x4*>(m + (row0 + int(l & 31u)) * stride + 8 * s + 4 * int(l >> 5));
}

// One layer with 64 outputs: out = act(W in), W's rows at w with `stride`
// halves per row, KSTEPS k-steps of input.  MASKED (backward pass): keep
// element i of fragment f where bit 4f + i of `mask` is set (relu_bits of the
// forward activation) instead of applying ReLU.
template <int KSTEPS, bool RELU, bool MASKED = false>
__device__ __forceinline__ void layer(const uint16_t* w, int stride, const bf16x4* in, bf16x4 out[8], uint32_t l,
                                      uint32_t mask = 0u) {
    f32x16 acc0 = {}, acc1 = {};
#pragma unroll
    for (int s = 0; s < KSTEPS; s++) {
        acc0 = mfma(load_a(w, stride, 0, s, l), in[s], acc0);
        acc1 = mfma(load_a(w, stride, 32, s, l), in[s], acc1);
    }
#pragma unroll
    for (int g = 0; g < 4; g++) {
        bf16x4 o0, o1;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float v0 = acc0[4 * g + i], v1 = acc1[4 * g + i];
            if (RELU) {
                v0 = v0 > 0.0f ? v0 : 0.0f;
                v1 = v1 > 0.0f ? v1 : 0.0f;
            }
            if (MASKED) {
                v0 = (mask >> (4 * g + i)) & 1u ? v0 : 0.0f;
                v1 = (mask >> (16 + 4 * g + i)) & 1u ? v1 : 0.0f;
            }
            o0[i] = bf(v0);
            o1[i] = bf(v1);
        }
        out[g] = o0;
        out[4 + g] = o1;
    }
}

// Output layer: z_j (j < 3) of the lane's sample in lanes 0..31.
__device__ __forceinline__ f32x16 output_layer(const uint16_t* w, const bf16x4 in[8], uint32_t l) {
    f32x16 acc = {};
#pragma unroll
    for (int s = 0; s < 8; s++) acc = mfma(load_a(w + MLP_FWD_W5, MLP_ROW, 0, s, l), in[s], acc);
    return acc;
}

// Bit 4f + i: element i of fragment f is positive (the ReLU derivative of a
// post-activation value).
__device__ __forceinline__ uint32_t relu_bits(const bf16x4 x[8]) {
    uint32_t m = 0u;
#pragma unroll
    for (int f = 0; f < 8; f++)
#pragma unroll
        for (int i = 0; i < 4; i++) m |= (x[f][i] > short(0) ? 1u : 0u) << (4 * f + i);
    return m;
}

__device__ __forceinline__ void load_record(const float4* records, uint32_t q, bool valid, MlpRecord& r) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float4 v = valid ? records[4 * size_t(q) + k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        r.v[4 * k + 0] = v.x;
        r.v[4 * k + 1] = v.y;
        r.v[4 * k + 2] = v.z;
        r.v[4 * k + 3] = v.w;
    }
}

__device__ __forceinline__ bf16x4 pack4(float v0, float v1, float v2, float v3) {
    bf16x4 v;
    v[0] = bf(v0);
    v[1] = bf(v1);
    v[2] = bf(v2);
    v[3] = bf(v3);
    return v;
}

// a_0 fragments: lane half h holds features 8s + 4h + 0..3, which are the
// encoding pairs 4s + 2h and 4s + 2h + 1, so each half encodes its own:
// mlp_encode_pair's arithmetic, with the pair's kind fixed by s (three steps
// per position axis) and the half entering only through selects.
__device__ __forceinline__ void encode_fragments(const MlpRecord& r, const MlpBounds& b, uint32_t l,
                                                 bf16x4 x[IN_STEPS]) {
    const bool hi = (l >> 5) != 0;
    const float u[3] = {mlp_unit(r, b, 0), mlp_unit(r, b, 1), mlp_unit(r, b, 2)};
#pragma unroll
    for (int s = 0; s < 9; s++) {
        // Frequencies i and i + 1 of axis s / 3.
        const int i = 4 * (s % 3) + (hi ? 2 : 0);
        float v0, v1, v2, v3;
        ttm::sincospi(u[s / 3] * ttm::exp2i(i), v0, v1);
        ttm::sincospi(u[s / 3] * ttm::exp2i(i + 1), v2, v3);
        x[s] = pack4(v0, v1, v2, v3);
    }
    // Features 72..79: normal, base colour r | base colour g, b, roughness,
    // metalness.
    x[9] = hi ? pack4(r.v[10], r.v[11], r.v[12], r.v[13]) : pack4(r.v[6], r.v[7], r.v[8], r.v[9]);
    // 80..87: [bounce == 1], 1, dir x | dir x (octave 1), dir y.
    float a0, a1, b0, b1;
    ttm::sincospi(r.v[3] * 2.0f, a0, a1);
    ttm::sincospi(hi ? r.v[4] : r.v[3], b0, b1);
    x[10] = pack4(hi ? a0 : mlp_bounce_is(r, 1), hi ? a1 : 1.0f, b0, b1);
    // 88..95: dir y (octave 1), dir z | dir z (octave 1), [bounce == 2],
    // [bounce >= 3].
    ttm::sincospi((hi ? r.v[5] : r.v[4]) * 2.0f, a0, a1);
    ttm::sincospi(r.v[5], b0, b1);
    x[11] = pack4(a0, a1, hi ? mlp_bounce_is(r, 2) : b0, hi ? (r.v[14] >= 3.0f ? 1.0f : 0.0f) : b1);
}

template <typename T>
__device__ __forceinline__ void copy_to_lds(T* dst, const T* src, int count) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
    int n = int(count * sizeof(T) / sizeof(uint4));
    for (int i = threadIdx.x; i < n; i += blockDim.x) d[i] = s[i];
}

// ------------------------------------------------------------ inference --
// min(*count_dev, count) records (count when count_dev is null).  out_rgb
// (if set) receives L; radiance (if set) gets radiance[path] +=
// throughput * L with (throughput, path) from scale_path.  Three waves per
// SIMD: what the LDS weights allow (three workgroups per CU).
__global__ __launch_bounds__(MLP_BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void mlp_infer_kernel(const uint16_t* wf, MlpBounds bounds,
                                                              const float4* records, const float4* scale_path,
                                                              const uint32_t* count_dev, uint32_t count_cap,
                                                              float* out_rgb, float4* radiance) {
    __shared__ __attribute__((aligned(16))) uint16_t w[MLP_FWD_HALVES];
    copy_to_lds(w, wf, MLP_FWD_HALVES);
    __syncthreads();
    const uint32_t count = count_dev ? min(*count_dev, count_cap) : count_cap;
    const uint32_t l = lane_id();
    const uint32_t waves = gridDim.x * (MLP_BLOCK / 64);
    for (uint32_t tile = blockIdx.x * (MLP_BLOCK / 64) + threadIdx.x / 64; tile * 32 < count; tile += waves) {
        // The weights stay in LDS: without this barrier the compiler hoists
        // every (loop-invariant) weight read out of the loop into registers,
        // which costs all but one wave per SIMD.
        __asm__ volatile("" ::: "memory");
        const uint32_t q = tile * 32 + (l & 31u);
        const bool valid = q < count;
        MlpRecord r;
        load_record(records, q, valid, r);
        bf16x4 a0[IN_STEPS], x[8], y[8];
        encode_fragments(r, bounds, l, a0);
        layer<IN_STEPS, true>(w, MLP_ROW0, a0, x, l);
        layer<8, true>(w + MLP_FWD_W1, MLP_ROW, x, y, l);
        layer<8, true>(w + MLP_FWD_W1 + 64 * MLP_ROW, MLP_ROW, y, x, l);
        layer<8, true>(w + MLP_FWD_W1 + 128 * MLP_ROW, MLP_ROW, x, y, l);
        layer<8, true>(w + MLP_FWD_W1 + 192 * MLP_ROW, MLP_ROW, y, x, l);
        f32x16 z = output_layer(w, x, l);
        if (l < 32 && valid) {
            float L0 = mlp_softplus(z[0]), L1 = mlp_softplus(z[1]), L2 = mlp_softplus(z[2]);
            if (out_rgb) {
                out_rgb[3 * size_t(q) + 0] = L0;
                out_rgb[3 * size_t(q) + 1] = L1;
                out_rgb[3 * size_t(q) + 2] = L2;
            }
            if (radiance) {
                float4 sp = scale_path[q];
                uint32_t p = __float_as_uint(sp.w);
                float4 acc = radiance[p];
                radiance[p] = make_float4(acc.x + sp.x * L0, acc.y + sp.y * L1, acc.z + sp.z * L2, acc.w);
            }
        }
    }
}

// Encoded features (tests).
__global__ void mlp_encode_kernel(MlpBounds bounds, const float4* records, uint32_t count, float* out) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    MlpRecord r;
    load_record(records, q, true, r);
    MlpFeatures f;
    mlp_encode(r, bounds, f);
    for (int k = 0; k < MLP_IN; k++) out[size_t(q) * MLP_IN + k] = f.f[k];
}

// Judge statistics over min(*count_dev, cap) records: stats[0] = count,
// then sums of predictions, targets, squared differences, squared targets
// and relative L2 terms (y - t)^2 / (y^2 + 0.01), over records and channels,
// in a fixed order.
__global__ __launch_bounds__(256) void mlp_judge_stats_kernel(const float* pred, const float4* targets,
                                                              const uint32_t* count_dev, uint32_t cap,
                                                              double* stats) {
    __shared__ double part[5][256];
    const uint32_t n = min(*count_dev, cap);
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
        float4 t4 = targets[i];
        const float t[3] = {t4.x, t4.y, t4.z};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double y = pred[3 * size_t(i) + c], tc = t[c], d = y - tc;
            s[0] += y;
            s[1] += tc;
            s[2] += d * d;
            s[3] += tc * tc;
            s[4] += d * d / (y * y + double(MLP_LOSS_EPS));
        }
    }
#pragma unroll
    for (int k = 0; k < 5; k++) part[k][threadIdx.x] = s[k];
    __syncthreads();
    for (uint32_t half = 128; half > 0; half >>= 1) {
        if (threadIdx.x < half)
#pragma unroll
            for (int k = 0; k < 5; k++) part[k][threadIdx.x] += part[k][threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        stats[0] = double(n);
#pragma unroll
        for (int k = 0; k < 5; k++) stats[1 + k] = part[k][0];
    }
}

// ------------------------------------------------------------- training --
// This step's records [first, first + n): consecutive slices of the
// min(*count_dev, cap) records (cap when count_dev is null).
__device__ __forceinline__ void step_slice(const uint32_t* count_dev, uint32_t cap, uint32_t step, uint32_t steps,
                                           uint32_t& first, uint32_t& n) {
    const uint32_t total = count_dev ? min(*count_dev, cap) : cap;
    first = uint32_t(uint64_t(total) * step / steps);
    n = uint32_t(uint64_t(total) * (step + 1) / steps) - first;
}

// Fragments x[0 .. NFRAG) of sample column q into rows of MLP_CHUNK halves.
// The compiler barrier after the stores keeps the next layer's weight reads
// from being hoisted above them (which would hold every layer's weights in
// registers at once and spill).
template <int NFRAG>
__device__ __forceinline__ void store_fragments(uint16_t* m, const bf16x4* x, uint32_t q, uint32_t l) {
    const int h = int(l >> 5);
#pragma unroll
    for (int s = 0; s < NFRAG; s++)
#pragma unroll
        for (int i = 0; i < 4; i++) m[size_t(8 * s + 4 * h + i) * MLP_CHUNK + q] = uint16_t(x[s][i]);
    __asm__ volatile("" ::: "memory");
}

__global__ __launch_bounds__(TRAIN_BLOCK) void mlp_train_kernel(const uint16_t* wf, const uint16_t* wb,
                                                                const uint16_t* wb5, MlpBounds bounds,
                                                                const float4* records, const float4* targets,
                                                                const uint32_t* count_dev, uint32_t cap,
                                                                uint32_t step, uint32_t steps, uint16_t* planes,
                                                                float* partial, float* loss_chunks) {
    uint32_t first, n;
    step_slice(count_dev, cap, step, steps, first, n);
    const uint32_t chunk = blockIdx.x;
    if (chunk * MLP_CHUNK >= n) return;  // the whole workgroup
    __shared__ __attribute__((aligned(16))) uint16_t w[MLP_FWD_HALVES];
    __shared__ __attribute__((aligned(16))) uint16_t wt[MLP_BWD_ROWS * MLP_ROW];
    __shared__ __attribute__((aligned(16))) uint16_t wt5[64 * MLP_BWD5_ROW];
    __shared__ float wave_loss[TRAIN_WAVES];
    copy_to_lds(w, wf, MLP_FWD_HALVES);
    copy_to_lds(wt, wb, MLP_BWD_ROWS * MLP_ROW);
    copy_to_lds(wt5, wb5, 64 * MLP_BWD5_ROW);
    __syncthreads();
    const uint32_t l = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    uint16_t* plane = planes + size_t(chunk) * PLANE_ROWS * MLP_CHUNK;

    // Forward and backward pass of the wave's tile.
    {
        const uint32_t ql = wave * 32 + (l & 31u);
        const uint32_t q = chunk * MLP_CHUNK + ql;
        const bool valid = q < n;
        MlpRecord r;
        load_record(records, first + q, valid, r);
        bf16x4 a0[IN_STEPS], a1[8], a2[8], a3[8], a4[8], a5[8];
        encode_fragments(r, bounds, l, a0);
        store_fragments<IN_STEPS>(plane + act_row(0) * MLP_CHUNK, a0, ql, l);
        // The backward pass needs only the sign pattern of a_1 .. a_5.
        layer<IN_STEPS, true>(w, MLP_ROW0, a0, a1, l);
        store_fragments<8>(plane + act_row(1) * MLP_CHUNK, a1, ql, l);
        const uint32_t m1 = relu_bits(a1);
        layer<8, true>(w + MLP_FWD_W1, MLP_ROW, a1, a2, l);
        store_fragments<8>(plane + act_row(2) * MLP_CHUNK, a2, ql, l);
        const uint32_t m2 = relu_bits(a2);
        layer<8, true>(w + MLP_FWD_W1 + 64 * MLP_ROW, MLP_ROW, a2, a3, l);
        store_fragments<8>(plane + act_row(3) * MLP_CHUNK, a3, ql, l);
        const uint32_t m3 = relu_bits(a3);
        layer<8, true>(w + MLP_FWD_W1 + 128 * MLP_ROW, MLP_ROW, a3, a4, l);
        store_fragments<8>(plane + act_row(4) * MLP_CHUNK, a4, ql, l);
        const uint32_t m4 = relu_bits(a4);
        layer<8, true>(w + MLP_FWD_W1 + 192 * MLP_ROW, MLP_ROW, a4, a5, l);
        store_fragments<8>(plane + act_row(5) * MLP_CHUNK, a5, ql, l);
        const uint32_t m5 = relu_bits(a5);
        f32x16 z = output_layer(w, a5, l);

        // Output delta (lanes 0..31 hold z_0..z_2 of their sample); samples
        // past the slice get a zero delta, hence zero gradient.
        const float inv_norm = 1.0f / (3.0f * float(n));
        bf16x4 dout = {0, 0, 0, 0};
        float loss = 0.0f;
        if (l < 32 && valid) {
            float4 t4 = targets[first + q];
            const float t[3] = {t4.x, t4.y, t4.z};
#pragma unroll
            for (int j = 0; j < 3; j++) {
                float y = mlp_softplus(z[j]);
                float diff = y - t[j];
                float den = y * y + MLP_LOSS_EPS;
                loss += diff * diff / den;
                float g = 2.0f * diff / den * inv_norm;
                dout[j] = bf(g * mlp_sigmoid(z[j]));
            }
        }
        if (l < 32)
#pragma unroll
            for (int j = 0; j < 3; j++) plane[size_t(delta_row(5) + j) * MLP_CHUNK + ql] = uint16_t(dout[j]);
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
        if (l == 0) wave_loss[wave] = loss;

        // d_4 = W_5^T d_5 masked by a_5 > 0.
        bf16x4 d[8], e[8];
        {
            f32x16 acc0 = {}, acc1 = {};
            acc0 = mfma(load_a(wt5, MLP_BWD5_ROW, 0, 0, l), dout, acc0);
            acc1 = mfma(load_a(wt5, MLP_BWD5_ROW, 32, 0, l), dout, acc1);
#pragma unroll
            for (int g = 0; g < 4; g++) {
                bf16x4 o0, o1;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    o0[i] = bf((m5 >> (4 * g + i)) & 1u ? acc0[4 * g + i] : 0.0f);
                    o1[i] = bf((m5 >> (16 + 4 * g + i)) & 1u ? acc1[4 * g + i] : 0.0f);
                }
                d[g] = o0;
                d[4 + g] = o1;
            }
        }
        store_fragments<8>(plane + delta_row(4) * MLP_CHUNK, d, ql, l);
        // d_{k-1} = W_k^T d_k masked by a_k > 0, k = 4 .. 1.
        layer<8, false, true>(wt + 3 * 64 * MLP_ROW, MLP_ROW, d, e, l, m4);
        store_fragments<8>(plane + delta_row(3) * MLP_CHUNK, e, ql, l);
        layer<8, false, true>(wt + 2 * 64 * MLP_ROW, MLP_ROW, e, d, l, m3);
        store_fragments<8>(plane + delta_row(2) * MLP_CHUNK, d, ql, l);
        layer<8, false, true>(wt + 1 * 64 * MLP_ROW, MLP_ROW, d, e, l, m2);
        store_fragments<8>(plane + delta_row(1) * MLP_CHUNK, e, ql, l);
        layer<8, false, true>(wt, MLP_ROW, e, d, l, m1);
        store_fragments<8>(plane + delta_row(0) * MLP_CHUNK, d, ql, l);
    }
    // The plane is complete (the barrier orders the workgroup's global
    // stores before the loads below).
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int k = 0; k < TRAIN_WAVES; k++) s += wave_loss[k];
        loss_chunks[chunk] = s;
    }

    // dW_k = d_k a_k^T over the chunk: 32 x 32 output tiles, the samples as
    // the k dimension.
    for (int id = int(wave); id < WGRAD_TILES; id += TRAIN_WAVES) {
        int k, t, u;
        if (id < 6) {
            k = 0;
            t = id / 3;
            u = id % 3;
        } else if (id < 22) {
            k = 1 + (id - 6) / 4;
            t = ((id - 6) % 4) / 2;
            u = (id - 6) % 2;
        } else {
            k = 5;
            t = 0;
            u = id - 22;
        }
        const uint16_t* dm = plane + size_t(delta_row(k)) * MLP_CHUNK;
        const uint16_t* am = plane + size_t(act_row(k)) * MLP_CHUNK;
        const uint32_t drow = 32 * t + (l & 31u), arow = 32 * u + (l & 31u);
        const bool dvalid = k < 5 || drow < 3u;
        f32x16 acc = {};
#pragma unroll 4
        for (int s = 0; s < MLP_CHUNK / 8; s++) {
            const int col = 8 * s + 4 * int(l >> 5);
            bf16x4 A = {0, 0, 0, 0};
            if (dvalid) A = *reinterpret_cast<const bf16x4*>(dm + size_t(drow) * MLP_CHUNK + col);
            bf16x4 B = *reinterpret_cast<const bf16x4*>(am + size_t(arow) * MLP_CHUNK + col);
            acc = mfma(A, B, acc);
        }
        const int in = k == 0 ? MLP_IN : 64;
        float* out = partial + size_t(chunk) * MLP_PARAMS + mlp_layer_offset(k);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const uint32_t row = 32 * t + 8 * (r / 4) + 4 * (l >> 5) + (r % 4);
            const uint32_t col = 32 * u + (l & 31u);
            if (k < 5 || row < 3u) out[row * in + col] = acc[r];
        }
    }
}

struct AdamArgs {
    float lr, beta1, beta2, eps, c1, c2;  // c1 = 1 / (1 - beta1^t), c2 = 1 / (1 - beta2^t)
};

// Layer, output row and input column of parameter idx.
__device__ __forceinline__ void param_position(uint32_t idx, uint32_t& k, uint32_t& o, uint32_t& i) {
    if (idx < uint32_t(MLP_W0)) {
        k = 0;
        o = idx / MLP_IN;
        i = idx % MLP_IN;
    } else if (idx < uint32_t(MLP_W5_OFFSET)) {
        uint32_t local = idx - MLP_W0;
        k = 1 + local / 4096u;
        local %= 4096u;
        o = local / 64u;
        i = local % 64u;
    } else {
        uint32_t local = idx - MLP_W5_OFFSET;
        k = 5;
        o = local / 64u;
        i = local % 64u;
    }
}

// Sums the step's chunk partials in chunk order; with apply, one Adam step
// that also writes the updated parameter's bf16 copies (the positions
// mlp_pack_kernel gives it; the copies' padding is written once, when the
// parameters are set).  A step with no records changes nothing.
__global__ void mlp_adam_kernel(float* master, float* m1, float* m2, const float* partial, const uint32_t* count_dev,
                                uint32_t cap, uint32_t step, uint32_t steps, AdamArgs a, int apply, float* grad_out,
                                uint16_t* wf, uint16_t* wb, uint16_t* wb5) {
    uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= uint32_t(MLP_PARAMS)) return;
    uint32_t first, n;
    step_slice(count_dev, cap, step, steps, first, n);
    if (n == 0) return;
    const uint32_t chunks = (n + MLP_CHUNK - 1) / MLP_CHUNK;
    float g = 0.0f;
    for (uint32_t c = 0; c < chunks; c++) g += partial[size_t(c) * MLP_PARAMS + idx];
    if (grad_out) grad_out[idx] = g;
    if (!apply) return;
    float mm = a.beta1 * m1[idx] + (1.0f - a.beta1) * g;
    float vv = a.beta2 * m2[idx] + (1.0f - a.beta2) * (g * g);
    m1[idx] = mm;
    m2[idx] = vv;
    const float w = master[idx] - a.lr * (mm * a.c1) / (ttm::sqrt(vv * a.c2) + a.eps);
    master[idx] = w;
    const uint16_t h = float_to_bf16(w);
    uint32_t k, o, i;
    param_position(idx, k, o, i);
    if (k == 0) {
        wf[o * MLP_ROW0 + i] = h;
    } else if (k < 5u) {
        wf[MLP_FWD_W1 + (64u * (k - 1u) + o) * MLP_ROW + i] = h;
        wb[(64u * (k - 1u) + i) * MLP_ROW + o] = h;
    } else {
        wf[MLP_FWD_W5 + o * MLP_ROW + i] = h;
        wb5[i * MLP_BWD5_ROW + o] = h;
    }
}

// bf16 copies: forward rows [out][in] (W_5 rows 3..31 zero), transposed
// W_1^T..W_4^T [in][out] and W_5^T [in][8]; padding zero.
__global__ void mlp_pack_kernel(const float* master, uint16_t* wf, uint16_t* wb, uint16_t* wb5) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t bwd = MLP_BWD_ROWS * MLP_ROW, bwd5 = 64 * MLP_BWD5_ROW;
    if (t < uint32_t(MLP_FWD_HALVES)) {
        float v = 0.0f;
        if (t < uint32_t(MLP_FWD_W1)) {
            uint32_t row = t / MLP_ROW0, col = t % MLP_ROW0;
            if (col < uint32_t(MLP_IN)) v = master[row * MLP_IN + col];
        } else if (t < uint32_t(MLP_FWD_W5)) {
            uint32_t local = t - MLP_FWD_W1;
            uint32_t row = local / MLP_ROW, col = local % MLP_ROW;
            uint32_t k = 1 + row / 64, o = row % 64;
            if (col < 64) v = master[mlp_layer_offset(int(k)) + o * 64 + col];
        } else {
            uint32_t local = t - MLP_FWD_W5;
            uint32_t row = local / MLP_ROW, col = local % MLP_ROW;
            if (row < 3 && col < 64) v = master[MLP_W5_OFFSET + row * 64 + col];
        }
        wf[t] = float_to_bf16(v);
        return;
    }
    t -= MLP_FWD_HALVES;
    if (t < bwd) {
        uint32_t row = t / MLP_ROW, col = t % MLP_ROW;  // W_k^T [in = row % 64][out = col]
        uint32_t k = 1 + row / 64, in = row % 64;
        float v = col < 64 ? master[mlp_layer_offset(int(k)) + col * 64 + in] : 0.0f;
        wb[t] = float_to_bf16(v);
        return;
    }
    t -= bwd;
    if (t < bwd5) {
        uint32_t in = t / MLP_BWD5_ROW, out = t % MLP_BWD5_ROW;
        wb5[t] = float_to_bf16(out < 3 ? master[MLP_W5_OFFSET + out * 64 + in] : 0.0f);
    }
}

// Training records in a fixed order.  Slots are taken in the order shading
// waves finish, so which records share a chunk, and with it the rounding of
// the gradient sums, would vary from run to run.  Each target's w holds a key
// (path x records a path takes + the path's record index; unique, below the
// span): records are scattered to an index by key, the occupied entries are
// counted by a scan, and gathered in key order, which makes training, and so
// the cached images, repeat bit for bit.  The gather leaves the index empty
// again for the next call.
constexpr uint32_t ORDER_EMPTY = 0xffffffffu;

__global__ __launch_bounds__(256) void mlp_order_fill_kernel(uint32_t* index, uint32_t span) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < span) index[j] = ORDER_EMPTY;
}

__global__ __launch_bounds__(256) void mlp_order_scatter_kernel(const float4* targets, const uint32_t* count_dev,
                                                                uint32_t cap, uint32_t span, uint32_t* index) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= min(*count_dev, cap)) return;
    const uint32_t key = __float_as_uint(targets[i].w);
    if (key < span) index[key] = i;
}

struct Occupied {
    __device__ uint32_t operator()(uint32_t v) const { return v != ORDER_EMPTY ? 1u : 0u; }
};

// Record index[j] goes to position pos[j]; the last thread writes the number
// of records gathered, which training then reads as its count.
__global__ __launch_bounds__(256) void mlp_order_gather_kernel(const float4* records, const float4* targets,
                                                               uint32_t* index, const uint32_t* pos, uint32_t span,
                                                               float4* rec_out, float4* tgt_out, uint32_t* count_out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= span) return;
    const uint32_t i = index[j];
    if (j == span - 1u) *count_out = pos[j] + (i != ORDER_EMPTY ? 1u : 0u);
    if (i == ORDER_EMPTY) return;
    const uint32_t o = pos[j];
#pragma unroll
    for (int k = 0; k < 4; k++) rec_out[4 * size_t(o) + k] = records[4 * size_t(i) + k];
    tgt_out[o] = targets[i];
    index[j] = ORDER_EMPTY;
}

uint32_t blocks_of(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

}  // namespace

// ---------------------------------------------------------------- host --
struct tt_mlp {
    tt_context* context = nullptr;
    float* master = nullptr;  // MLP_PARAMS
    float* m1 = nullptr;
    float* m2 = nullptr;
    uint16_t* wf = nullptr;   // forward bf16
    uint16_t* wb = nullptr;   // transposed bf16
    uint16_t* wb5 = nullptr;
    MlpBounds bounds{{0.0f, 0.0f, 0.0f}, {1.0f, 1.0f, 1.0f}};
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-8f;
    uint64_t step = 0;
    // Training runs on the context's API stream after `ready`; `trained`
    // marks the end of the last training enqueued.
    hipEvent_t ready = nullptr;
    hipEvent_t trained = nullptr;
    bool trained_recorded = false;
    tt::EventPair train_time;  // around the last timed training (created on first use)
    // Training scratch for `chunk_capacity` chunks a step, grown on demand:
    // staging planes, fp32 partials, and MAX_STEPS rows of chunk losses.
    uint32_t chunk_capacity = 0;
    void* train_memory = nullptr;
    uint16_t* planes = nullptr;
    float* partial = nullptr;
    float* loss = nullptr;
    // Record ordering for `order_capacity` records over keys below
    // `order_span`, grown on demand: the index by key (empty between calls),
    // the scanned positions, the ordered records and targets, their count,
    // and the scan's temporary storage.
    uint32_t order_capacity = 0, order_span = 0;
    void* order_memory = nullptr;
    uint32_t* order_index = nullptr;
    uint32_t* order_pos = nullptr;
    float4* order_rec = nullptr;
    float4* order_tgt = nullptr;
    uint32_t* order_count = nullptr;
    void* order_temp = nullptr;
    size_t order_temp_bytes = 0;
};

namespace tt {

using tt::fail;

static hipStream_t train_stream(tt_mlp* m) { return m->context->stream; }

tt_status mlp_repack(tt_mlp* m, hipStream_t s) {
    tt_context* c = m->context;
    uint32_t total = MLP_FWD_HALVES + MLP_BWD_ROWS * MLP_ROW + 64 * MLP_BWD5_ROW;
    mlp_pack_kernel<<<blocks_of(total, 256), 256, 0, s>>>(m->master, m->wf, m->wb, m->wb5);
    TT_HIP(c, hipGetLastError());
    return TT_OK;
}

static tt_status ensure_train(tt_mlp* m, uint32_t chunks) {
    tt_context* c = m->context;
    if (chunks <= m->chunk_capacity) return TT_OK;
    if (m->train_memory) {
        TT_HIP(c, hipStreamSynchronize(train_stream(m)));
        TT_HIP(c, hipFree(m->train_memory));
        m->train_memory = nullptr;
        m->chunk_capacity = 0;
    }
    size_t planes = size_t(chunks) * PLANE_ROWS * MLP_CHUNK * sizeof(uint16_t);
    size_t partial = size_t(chunks) * MLP_PARAMS * sizeof(float);
    size_t loss = size_t(MAX_STEPS) * chunks * sizeof(float);
    TT_HIP(c, hipMalloc(&m->train_memory, planes + partial + loss));
    char* p = static_cast<char*>(m->train_memory);
    m->planes = reinterpret_cast<uint16_t*>(p);
    m->partial = reinterpret_cast<float*>(p + planes);
    m->loss = reinterpret_cast<float*>(p + planes + partial);
    m->chunk_capacity = chunks;
    return TT_OK;
}

static tt_status ensure_order(tt_mlp* m, uint32_t cap, uint32_t span) {
    tt_context* c = m->context;
    if (cap <= m->order_capacity && span <= m->order_span) return TT_OK;
    hipStream_t s = train_stream(m);
    if (m->order_memory) {
        TT_HIP(c, hipStreamSynchronize(s));
        TT_HIP(c, hipFree(m->order_memory));
        m->order_memory = nullptr;
        m->order_capacity = m->order_span = 0;
    }
    cap = std::max(cap, m->order_capacity);
    span = std::max(span, m->order_span);
    size_t temp = 0;
    {
        uint32_t* u = nullptr;
        hipError_t e = rocprim::exclusive_scan(nullptr, temp, rocprim::make_transform_iterator(u, Occupied()), u, 0u,
                                               size_t(span), rocprim::plus<uint32_t>(), s);
        if (e != hipSuccess) return tt::hip_fail(c, e, "record order scan size");
    }
    auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t index = align(size_t(span) * sizeof(uint32_t));
    const size_t rec = align(4 * size_t(cap) * sizeof(float4)), tgt = align(size_t(cap) * sizeof(float4));
    TT_HIP(c, hipMalloc(&m->order_memory, 2 * index + rec + tgt + 256 + align(temp)));
    char* p = static_cast<char*>(m->order_memory);
    m->order_index = reinterpret_cast<uint32_t*>(p);
    m->order_pos = reinterpret_cast<uint32_t*>(p + index);
    m->order_rec = reinterpret_cast<float4*>(p + 2 * index);
    m->order_tgt = reinterpret_cast<float4*>(p + 2 * index + rec);
    m->order_count = reinterpret_cast<uint32_t*>(p + 2 * index + rec + tgt);
    m->order_temp = p + 2 * index + rec + tgt + 256;
    m->order_temp_bytes = temp;
    mlp_order_fill_kernel<<<blocks_of(span, 256), 256, 0, s>>>(m->order_index, span);
    TT_HIP(c, hipGetLastError());
    m->order_capacity = cap;
    m->order_span = span;
    return TT_OK;
}

tt_context* mlp_context(const tt_mlp* m) { return m->context; }

tt_status mlp_set_scene_bounds(tt_mlp* m, const float lo[3], const float hi[3]) {
    for (int k = 0; k < 3; k++) {
        float ext = hi[k] - lo[k];
        m->bounds.lo[k] = lo[k];
        m->bounds.inv_extent[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
    return TT_OK;
}

tt_status mlp_wait_trained(tt_mlp* m, hipStream_t stream) {
    if (m->trained_recorded && stream != train_stream(m)) TT_HIP(m->context, hipStreamWaitEvent(stream, m->trained, 0));
    return TT_OK;
}

tt_status mlp_infer_device(tt_mlp* m, hipStream_t stream, const float4* records, const float4* scale_path,
                           const uint32_t* count_dev, uint32_t count, float* out_rgb, float4* radiance) {
    tt_context* c = m->context;
    if (count == 0) return TT_OK;
    tt_status st = mlp_wait_trained(m, stream);
    if (st != TT_OK) return st;
    uint32_t tiles = blocks_of(count, 32);
    uint32_t grid = std::min<uint32_t>(blocks_of(tiles, MLP_BLOCK / 64), uint32_t(c->cu_count) * 2u);
    mlp_infer_kernel<<<grid, MLP_BLOCK, 0, stream>>>(m->wf, m->bounds, records, scale_path, count_dev, count, out_rgb,
                                                     radiance);
    TT_HIP(c, hipGetLastError());
    return TT_OK;
}

tt_status mlp_judge_device(tt_mlp* m, hipStream_t stream, const float4* records, const float4* targets,
                           const uint32_t* count_dev, uint32_t capacity, float* pred, double* stats) {
    tt_context* c = m->context;
    tt_status st = mlp_infer_device(m, stream, records, nullptr, count_dev, capacity, pred, nullptr);
    if (st != TT_OK) return st;
    mlp_judge_stats_kernel<<<1, 256, 0, stream>>>(pred, targets, count_dev, capacity, stats);
    TT_HIP(c, hipGetLastError());
    return TT_OK;
}

tt_status mlp_train_device(tt_mlp* m, hipStream_t after, const float4* records, const float4* targets,
                           const uint32_t* count_dev, uint32_t capacity, uint32_t steps, bool apply, float* grad_out,
                           bool timed, uint32_t order_span) {
    tt_context* c = m->context;
    if (capacity == 0 || steps == 0) return TT_OK;
    if (steps > MAX_STEPS) return fail(c, TT_INVALID_ARGUMENT, "more than 64 training steps");
    if (order_span && !count_dev) return fail(c, TT_INVALID_ARGUMENT, "ordered training needs the record count");
    // The largest slice has ceil(capacity / steps) records.
    const uint32_t chunks = blocks_of(blocks_of(capacity, steps), MLP_CHUNK);
    tt_status st = ensure_train(m, chunks);
    if (st != TT_OK) return st;
    if (order_span) {
        st = ensure_order(m, capacity, order_span);
        if (st != TT_OK) return st;
    }
    hipStream_t s = train_stream(m);
    if (after != s) {
        TT_HIP(c, hipEventRecord(m->ready, after));
        TT_HIP(c, hipStreamWaitEvent(s, m->ready, 0));
    }
    if (timed) {
        if (!m->train_time.begin) TT_HIP(c, hipEventCreate(&m->train_time.begin));
        if (!m->train_time.end) TT_HIP(c, hipEventCreate(&m->train_time.end));
        TT_HIP(c, hipEventRecord(m->train_time.begin, s));
    }
    if (order_span) {
        mlp_order_scatter_kernel<<<blocks_of(capacity, 256), 256, 0, s>>>(targets, count_dev, capacity, order_span,
                                                                          m->order_index);
        TT_HIP(c, hipGetLastError());
        size_t temp = m->order_temp_bytes;
        TT_HIP(c, rocprim::exclusive_scan(m->order_temp, temp, rocprim::make_transform_iterator(m->order_index,
                                                                                                Occupied()),
                                          m->order_pos, 0u, size_t(order_span), rocprim::plus<uint32_t>(), s));
        mlp_order_gather_kernel<<<blocks_of(order_span, 256), 256, 0, s>>>(records, targets, m->order_index,
                                                                           m->order_pos, order_span, m->order_rec,
                                                                           m->order_tgt, m->order_count);
        TT_HIP(c, hipGetLastError());
        records = m->order_rec;
        targets = m->order_tgt;
        count_dev = m->order_count;
    }
    for (uint32_t i = 0; i < steps; i++) {
        mlp_train_kernel<<<chunks, TRAIN_BLOCK, 0, s>>>(m->wf, m->wb, m->wb5, m->bounds, records, targets, count_dev,
                                                        capacity, i, steps, m->planes, m->partial,
                                                        m->loss + size_t(i) * m->chunk_capacity);
        TT_HIP(c, hipGetLastError());
        AdamArgs a{m->lr, m->beta1, m->beta2, m->eps, 1.0f, 1.0f};
        if (apply) {
            m->step++;
            a.c1 = float(1.0 / (1.0 - std::pow(double(m->beta1), double(m->step))));
            a.c2 = float(1.0 / (1.0 - std::pow(double(m->beta2), double(m->step))));
        }
        mlp_adam_kernel<<<blocks_of(MLP_PARAMS, 256), 256, 0, s>>>(m->master, m->m1, m->m2, m->partial, count_dev,
                                                                   capacity, i, steps, a, apply ? 1 : 0, grad_out,
                                                                   m->wf, m->wb, m->wb5);
        TT_HIP(c, hipGetLastError());
    }
    if (timed) TT_HIP(c, hipEventRecord(m->train_time.end, s));
    TT_HIP(c, hipEventRecord(m->trained, s));
    m->trained_recorded = true;
    return TT_OK;
}

tt_status mlp_train_ms(tt_mlp* m, double* out) {
    tt_context* c = m->context;
    *out = 0.0;
    if (!m->train_time.end) return TT_OK;
    TT_HIP(c, hipEventSynchronize(m->train_time.end));
    float ms = 0.0f;
    TT_HIP(c, hipEventElapsedTime(&ms, m->train_time.begin, m->train_time.end));
    *out = ms;
    return TT_OK;
}

tt_status mlp_first_step_loss(tt_mlp* m, uint32_t count, uint32_t steps, double* out) {
    tt_context* c = m->context;
    *out = 0.0;
    if (steps == 0 || m->chunk_capacity == 0) return TT_OK;
    const uint32_t n = count / steps;  // the first slice
    if (n == 0) return TT_OK;
    std::vector<float> v(blocks_of(n, MLP_CHUNK));
    TT_HIP(c, hipMemcpyAsync(v.data(), m->loss, v.size() * sizeof(float), hipMemcpyDeviceToHost, train_stream(m)));
    TT_HIP(c, hipStreamSynchronize(train_stream(m)));
    double s = 0.0;
    for (float x : v) s += x;
    *out = s / (3.0 * n);
    return TT_OK;
}

}  // namespace tt

namespace {

// Host PCG32 (default stream) for the weight initialisation.
struct HostPcg {
    uint64_t s;
    explicit HostPcg(uint64_t seed) {
        const uint64_t inc = 0xda3e39cb94b95bdbull;
        s = inc;
        s += seed;
        s = s * 6364136223846793005ull + inc;
    }
    uint32_t next() {
        uint64_t old = s;
        s = old * 6364136223846793005ull + 0xda3e39cb94b95bdbull;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    float uniform() { return float(next() >> 8) * (1.0f / 16777216.0f); }
};

// Host operations of a cache run on the API stream, after every render and
// training enqueued so far.
hipStream_t host_stream(tt_mlp* m) { return tt::api_stream(m->context); }

}  // namespace

extern "C" {

tt_status tt_mlp_create(tt_context* c, const tt_mlp_desc* desc, tt_mlp** out) try {
    if (!c || !out) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    tt_mlp* m = new (std::nothrow) tt_mlp();
    if (!m) return fail(c, TT_OUT_OF_MEMORY, "mlp allocation");
    m->context = c;
    tt::context_retain(c);
    if (desc && desc->learning_rate > 0.0f) m->lr = desc->learning_rate;
    auto bail = [&](tt_status s) {
        tt_mlp_destroy(m);
        return s;
    };
    TT_HIP(c, hipSetDevice(c->device));
    if (hipEventCreateWithFlags(&m->ready, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->trained, hipEventDisableTiming) != hipSuccess)
        return bail(TT_HIP_ERROR);
    size_t bwd = size_t(MLP_BWD_ROWS) * MLP_ROW, bwd5 = 64 * MLP_BWD5_ROW;
    if (hipMalloc(&m->master, 3 * MLP_PARAMS * sizeof(float)) != hipSuccess) return bail(TT_OUT_OF_MEMORY);
    m->m1 = m->master + MLP_PARAMS;
    m->m2 = m->master + 2 * MLP_PARAMS;
    if (hipMalloc(&m->wf, (MLP_FWD_HALVES + bwd + bwd5) * sizeof(uint16_t)) != hipSuccess)
        return bail(TT_OUT_OF_MEMORY);
    m->wb = m->wf + MLP_FWD_HALVES;
    m->wb5 = m->wb + bwd;
    // He-uniform initialisation, U(-sqrt(6 / fan_in), sqrt(6 / fan_in));
    // Adam state 0.
    std::vector<float> host(3 * MLP_PARAMS, 0.0f);
    HostPcg rng(desc ? desc->seed : 0);
    for (int i = 0; i < MLP_PARAMS; i++) {
        const float fan_in = i < MLP_W0 ? float(MLP_IN) : 64.0f;
        host[i] = (2.0f * rng.uniform() - 1.0f) * std::sqrt(6.0f / fan_in);
    }
    hipStream_t s = host_stream(m);
    TT_HIP(c, hipMemcpyAsync(m->master, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, s));
    tt_status st = tt::mlp_repack(m, s);
    if (st != TT_OK) return bail(st);
    TT_HIP(c, hipStreamSynchronize(s));
    *out = m;
    return TT_OK;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

void tt_mlp_destroy(tt_mlp* m) try {
    if (!m) return;
    tt_context* c = m->context;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(tt::api_stream(c));
    if (m->master) (void)hipFree(m->master);
    if (m->wf) (void)hipFree(m->wf);
    if (m->train_memory) (void)hipFree(m->train_memory);
    if (m->order_memory) (void)hipFree(m->order_memory);
    for (hipEvent_t e : {m->ready, m->trained, m->train_time.begin, m->train_time.end})
        if (e) (void)hipEventDestroy(e);
    delete m;
    tt::context_release(c);
} catch (...) {
}

tt_status tt_mlp_set_bounds(tt_mlp* m, const float lo[3], const float hi[3]) try {
    if (!m || !lo || !hi) return fail(m ? m->context : nullptr, TT_INVALID_ARGUMENT, "null argument");
    return tt::mlp_set_scene_bounds(m, lo, hi);
} catch (...) {
    return tt::exception_status(m && m->context ? &m->context->error : nullptr);
}

tt_status tt_mlp_params(tt_mlp* m, float* get, const float* set, uint32_t count) try {
    if (!m) return TT_INVALID_ARGUMENT;
    tt_context* c = m->context;
    if (count != uint32_t(MLP_PARAMS)) return fail(c, TT_INVALID_ARGUMENT, "parameter count is 22720");
    TT_HIP(c, hipSetDevice(c->device));
    hipStream_t s = host_stream(m);
    if (get) {
        TT_HIP(c, hipMemcpyAsync(get, m->master, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, s));
        TT_HIP(c, hipStreamSynchronize(s));
    }
    if (set) {
        TT_HIP(c, hipMemcpyAsync(m->master, set, MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, s));
        tt_status st = tt::mlp_repack(m, s);
        if (st != TT_OK) return st;
        TT_HIP(c, hipStreamSynchronize(s));
    }
    return TT_OK;
} catch (...) {
    return tt::exception_status(m && m->context ? &m->context->error : nullptr);
}

// Host-array entry points (tests): op 0 encode (out: 96 floats per record),
// 1 infer (out: 3 floats), 2 gradient (targets: 4 floats per record; out:
// MLP_PARAMS floats; loss), 3 one Adam step (targets; loss).
tt_status tt_mlp_run(tt_mlp* m, int op, const float* records, const float* targets, uint32_t n, float* out,
                     double* loss) try {
    if (!m || !records || n == 0) return fail(m ? m->context : nullptr, TT_INVALID_ARGUMENT, "null argument");
    tt_context* c = m->context;
    if ((op >= 2 && !targets) || (op <= 2 && !out) || op < 0 || op > 3)
        return fail(c, TT_INVALID_ARGUMENT, "bad tt_mlp_run arguments");
    TT_HIP(c, hipSetDevice(c->device));
    hipStream_t s = host_stream(m);
    size_t rec_bytes = size_t(n) * MLP_RECORD_FLOATS * sizeof(float);
    size_t tgt_bytes = size_t(n) * 4 * sizeof(float);
    size_t out_floats = op == 0 ? size_t(n) * MLP_IN : (op == 1 ? size_t(n) * 3 : (op == 2 ? MLP_PARAMS : 0));
    char* d = nullptr;
    TT_HIP(c, hipMalloc(&d, rec_bytes + tgt_bytes + out_floats * sizeof(float) + 16));
    float4* drec = reinterpret_cast<float4*>(d);
    float4* dtgt = reinterpret_cast<float4*>(d + rec_bytes);
    float* dout = reinterpret_cast<float*>(d + rec_bytes + tgt_bytes);
    tt_status st = TT_OK;
    hipError_t e = hipMemcpyAsync(drec, records, rec_bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && targets) e = hipMemcpyAsync(dtgt, targets, tgt_bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) st = tt::hip_fail(c, e, "tt_mlp_run upload");
    if (st == TT_OK) {
        if (op == 0) {
            mlp_encode_kernel<<<blocks_of(n, 64), 64, 0, s>>>(m->bounds, drec, n, dout);
            e = hipGetLastError();
            if (e != hipSuccess) st = tt::hip_fail(c, e, "encode");
        } else if (op == 1) {
            st = tt::mlp_infer_device(m, s, drec, nullptr, nullptr, n, dout, nullptr);
        } else {
            st = tt::mlp_train_device(m, s, drec, dtgt, nullptr, n, 1, op == 3, op == 2 ? dout : nullptr);
        }
    }
    if (st == TT_OK && out_floats) {
        e = hipMemcpyAsync(out, dout, out_floats * sizeof(float), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) st = tt::hip_fail(c, e, "tt_mlp_run download");
    }
    if (st == TT_OK && op >= 2 && loss) st = tt::mlp_first_step_loss(m, n, 1, loss);
    e = hipStreamSynchronize(s);
    if (st == TT_OK && e != hipSuccess) st = tt::hip_fail(c, e, "tt_mlp_run");
    (void)hipFree(d);
    return st;
} catch (...) {
    return tt::exception_status(m && m->context ? &m->context->error : nullptr);
}

}  // extern "C"
