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
// LDS as bf16 rows of 68 (conflict-free for the 32 rows a wave reads).
//
// Training is one Adam step per call over a batch:
//   mlp_train_kernel   forward + backward per 32-sample tile; stores the
//                      activations a_0..a_5 and the deltas of every layer
//                      (bf16, [feature][sample]) and per-tile loss sums
//   mlp_wgrad_kernel   dW_l = delta_l a_l^T over 256-sample chunks (MFMA)
//                      into per-chunk partials
//   mlp_adam_kernel    sums the partials in chunk order (deterministic) and
//                      applies Adam to the fp32 master weights
//   mlp_pack_kernel    refreshes the bf16 forward / transposed copies
#include <hip/hip_runtime.h>

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

constexpr int MLP_BLOCK = 256;  // 4 waves
constexpr int MLP_CHUNK = 256;  // samples per weight-gradient partial
constexpr int W5_ROW0 = 5 * 64;  // first forward row of W_5

__device__ __forceinline__ f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ short bf(float x) { return short(float_to_bf16(x)); }
__device__ __forceinline__ float unbf(short x) { return bf16_to_float(uint16_t(x)); }

// A operand: rows row0 + (l % 32) of a row-major bf16 matrix with `stride`
// halves per row, k-step s.
__device__ __forceinline__ bf16x4 load_a(const uint16_t* m, int stride, int row0, int s, uint32_t l) {
    return *reinterpret_cast<const bf16x4*>(m + (row0 + int(l & 31u)) * stride + 8 * s + 4 * int(l >> 5));
}

// One 64-wide layer: out = act(W[row0 .. row0+63] in), W in LDS rows of
// MLP_ROW.  RELU_MASK: multiply by [mask > 0] instead of applying ReLU to
// the result (backward pass).
template <bool RELU>
__device__ __forceinline__ void layer64(const uint16_t* w, int row0, const bf16x4 in[8], bf16x4 out[8], uint32_t l,
                                        const bf16x4* mask = nullptr) {
    f32x16 acc0 = {}, acc1 = {};
#pragma unroll
    for (int s = 0; s < 8; s++) {
        acc0 = mfma(load_a(w, MLP_ROW, row0, s, l), in[s], acc0);
        acc1 = mfma(load_a(w, MLP_ROW, row0 + 32, s, l), in[s], acc1);
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
            if (mask) {
                v0 = unbf(mask[g][i]) > 0.0f ? v0 : 0.0f;
                v1 = unbf(mask[4 + g][i]) > 0.0f ? v1 : 0.0f;
            }
            o0[i] = bf(v0);
            o1[i] = bf(v1);
        }
        out[g] = o0;
        out[4 + g] = o1;
    }
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

// a_0 fragments: lane half h holds features 8s + 4h + 0..3.
__device__ __forceinline__ void encode_fragments(const MlpRecord& r, const MlpBounds& b, uint32_t l, bf16x4 x[8]) {
    MlpFeatures f;
    mlp_encode(r, b, f);
    const bool hi = (l >> 5) != 0;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        bf16x4 v;
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = bf(hi ? f.f[8 * s + 4 + i] : f.f[8 * s + i]);
        x[s] = v;
    }
}

// Output layer: z_j (j < 3) of the lane's sample in lanes 0..31.
__device__ __forceinline__ f32x16 output_layer(const uint16_t* w, const bf16x4 in[8], uint32_t l) {
    f32x16 acc = {};
#pragma unroll
    for (int s = 0; s < 8; s++) acc = mfma(load_a(w, MLP_ROW, W5_ROW0, s, l), in[s], acc);
    return acc;
}

template <typename T>
__device__ __forceinline__ void copy_to_lds(T* dst, const T* src, int count) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
    int n = int(count * sizeof(T) / sizeof(uint4));
    for (int i = threadIdx.x; i < n; i += blockDim.x) d[i] = s[i];
}

// ------------------------------------------------------------ inference --
// out_rgb (if set) receives L; radiance (if set) gets radiance[path] +=
// throughput * L with (throughput, path) from scale_path.
__global__ __launch_bounds__(MLP_BLOCK) void mlp_infer_kernel(const uint16_t* wf, MlpBounds bounds,
                                                              const float4* records, const float4* scale_path,
                                                              const uint32_t* count_dev, uint32_t count_host,
                                                              float* out_rgb, float4* radiance) {
    __shared__ __attribute__((aligned(16))) uint16_t w[MLP_FWD_ROWS * MLP_ROW];
    copy_to_lds(w, wf, MLP_FWD_ROWS * MLP_ROW);
    __syncthreads();
    const uint32_t count = count_dev ? *count_dev : count_host;
    const uint32_t l = lane_id();
    const uint32_t waves = gridDim.x * (MLP_BLOCK / 64);
    for (uint32_t tile = blockIdx.x * (MLP_BLOCK / 64) + threadIdx.x / 64; tile * 32 < count; tile += waves) {
        const uint32_t q = tile * 32 + (l & 31u);
        const bool valid = q < count;
        MlpRecord r;
        load_record(records, q, valid, r);
        bf16x4 a[8], b[8];
        encode_fragments(r, bounds, l, a);
        layer64<true>(w, 0, a, b, l);
        layer64<true>(w, 64, b, a, l);
        layer64<true>(w, 128, a, b, l);
        layer64<true>(w, 192, b, a, l);
        layer64<true>(w, 256, a, b, l);
        f32x16 z = output_layer(w, b, l);
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

// Encoded features (debugging / tests).
__global__ void mlp_encode_kernel(MlpBounds bounds, const float4* records, uint32_t count, float* out) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    MlpRecord r;
    load_record(records, q, true, r);
    MlpFeatures f;
    mlp_encode(r, bounds, f);
    for (int k = 0; k < MLP_IN; k++) out[size_t(q) * MLP_IN + k] = f.f[k];
}

// ------------------------------------------------------------- training --
// acts: a_0..a_5, deltas: d_0..d_5 (d_l = dLoss/dz of W_l's output; d_5 has
// 3 valid rows), each [64 rows][nstride samples] bf16.
__device__ __forceinline__ void store_fragments(uint16_t* m, size_t nstride, const bf16x4 x[8], uint32_t q,
                                                uint32_t l) {
    const int h = int(l >> 5);
#pragma unroll
    for (int s = 0; s < 8; s++)
#pragma unroll
        for (int i = 0; i < 4; i++) m[size_t(8 * s + 4 * h + i) * nstride + q] = uint16_t(x[s][i]);
}

__global__ __launch_bounds__(MLP_BLOCK) void mlp_train_kernel(const uint16_t* wf, const uint16_t* wb,
                                                              const uint16_t* wb5, MlpBounds bounds,
                                                              const float4* records, const float4* targets,
                                                              uint32_t n, uint32_t nstride, float inv_norm,
                                                              uint16_t* acts, uint16_t* deltas, float* loss_tiles) {
    __shared__ __attribute__((aligned(16))) uint16_t w[MLP_FWD_ROWS * MLP_ROW];
    __shared__ __attribute__((aligned(16))) uint16_t wt[MLP_BWD_ROWS * MLP_ROW];
    __shared__ __attribute__((aligned(16))) uint16_t wt5[64 * MLP_BWD5_ROW];
    copy_to_lds(w, wf, MLP_FWD_ROWS * MLP_ROW);
    copy_to_lds(wt, wb, MLP_BWD_ROWS * MLP_ROW);
    copy_to_lds(wt5, wb5, 64 * MLP_BWD5_ROW);
    __syncthreads();
    const uint32_t l = lane_id();
    const uint32_t waves = gridDim.x * (MLP_BLOCK / 64);
    const size_t plane = size_t(64) * nstride;
    for (uint32_t tile = blockIdx.x * (MLP_BLOCK / 64) + threadIdx.x / 64; tile * 32 < nstride; tile += waves) {
        const uint32_t q = tile * 32 + (l & 31u);
        const bool valid = q < n;
        MlpRecord r;
        load_record(records, q, valid, r);
        bf16x4 a[6][8];
        encode_fragments(r, bounds, l, a[0]);
#pragma unroll
        for (int k = 0; k < 5; k++) layer64<true>(w, 64 * k, a[k], a[k + 1], l);
        f32x16 z = output_layer(w, a[5], l);
#pragma unroll
        for (int k = 0; k < 6; k++) store_fragments(acts + k * plane, nstride, a[k], q, l);

        // Output delta (lanes 0..31 hold z_0..z_2 of their sample).
        bf16x4 dout = {0, 0, 0, 0};
        float loss = 0.0f;
        if (l < 32 && valid) {
            float4 t4 = targets[q];
            float t[3] = {t4.x, t4.y, t4.z};
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
        if (l < 32) {
#pragma unroll
            for (int j = 0; j < 3; j++) deltas[5 * plane + size_t(j) * nstride + q] = uint16_t(dout[j]);
        }
        // Tile loss (sum over the wave's lanes) for monitoring.
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
        if (l == 0) loss_tiles[tile] = loss;

        // delta_4 (output of W_4) = W_5^T dout masked by a_5 > 0.
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
                    o0[i] = bf(unbf(a[5][g][i]) > 0.0f ? acc0[4 * g + i] : 0.0f);
                    o1[i] = bf(unbf(a[5][4 + g][i]) > 0.0f ? acc1[4 * g + i] : 0.0f);
                }
                d[g] = o0;
                d[4 + g] = o1;
            }
        }
        store_fragments(deltas + 4 * plane, nstride, d, q, l);
        // delta_{k-1} = W_k^T delta_k masked by a_k > 0, k = 4 .. 1.
        layer64<false>(wt, 3 * 64, d, e, l, a[4]);
        store_fragments(deltas + 3 * plane, nstride, e, q, l);
        layer64<false>(wt, 2 * 64, e, d, l, a[3]);
        store_fragments(deltas + 2 * plane, nstride, d, q, l);
        layer64<false>(wt, 1 * 64, d, e, l, a[2]);
        store_fragments(deltas + 1 * plane, nstride, e, q, l);
        layer64<false>(wt, 0, e, d, l, a[1]);
        store_fragments(deltas + 0 * plane, nstride, d, q, l);
    }
}

// dW_k [out][in] over samples [chunk * MLP_CHUNK, +MLP_CHUNK): one wave per
// (chunk, layer); partial[k][chunk][out][in] (fp32, out rows 0..63).
__global__ __launch_bounds__(64) void mlp_wgrad_kernel(const uint16_t* acts, const uint16_t* deltas, uint32_t nstride,
                                                       float* partial) {
    const uint32_t chunk = blockIdx.x, k = blockIdx.y;
    const uint32_t chunks = gridDim.x;
    const uint32_t l = threadIdx.x;
    const size_t plane = size_t(64) * nstride;
    const uint16_t* dm = deltas + k * plane;
    const uint16_t* am = acts + k * plane;
    const int tiles_out = k == 5 ? 1 : 2;
    f32x16 acc[2][2] = {};
    const uint32_t q0 = chunk * MLP_CHUNK;
    for (int s = 0; s < MLP_CHUNK / 8; s++) {
        const size_t col = q0 + 8 * s + 4 * (l >> 5);
        bf16x4 b0 = *reinterpret_cast<const bf16x4*>(am + size_t(l & 31u) * nstride + col);
        bf16x4 b1 = *reinterpret_cast<const bf16x4*>(am + size_t(32 + (l & 31u)) * nstride + col);
#pragma unroll
        for (int t = 0; t < 2; t++) {
            if (t < tiles_out) {
                uint32_t row = 32 * t + (l & 31u);
                bf16x4 a = {0, 0, 0, 0};
                if (k != 5 || row < 3) a = *reinterpret_cast<const bf16x4*>(dm + size_t(row) * nstride + col);
                acc[t][0] = mfma(a, b0, acc[t][0]);
                acc[t][1] = mfma(a, b1, acc[t][1]);
            }
        }
    }
    float* out = partial + (size_t(k) * chunks + chunk) * 64 * 64;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        if (t >= tiles_out) continue;
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                uint32_t row = 32 * t + 8 * (r / 4) + 4 * (l >> 5) + (r % 4);
                uint32_t col = 32 * u + (l & 31u);
                out[row * 64 + col] = acc[t][u][r];
            }
    }
}

struct AdamArgs {
    float lr, beta1, beta2, eps, c1, c2;  // c1 = 1 / (1 - beta1^t), c2 = 1 / (1 - beta2^t)
};

__device__ __forceinline__ void param_position(uint32_t idx, uint32_t& k, uint32_t& o, uint32_t& i) {
    k = idx < 5u * 4096u ? idx / 4096u : 5u;
    uint32_t local = idx - k * 4096u;
    o = local / 64u;
    i = local % 64u;
}

// With apply, the step also refreshes the bf16 copies of the updated
// parameter (the positions mlp_pack_kernel gives it; the copies' padding is
// written once, when the parameters are set).
__global__ void mlp_adam_kernel(float* master, float* m1, float* m2, const float* partial, uint32_t chunks,
                                AdamArgs a, int apply, float* grad_out, uint16_t* wf, uint16_t* wb, uint16_t* wb5) {
    uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= uint32_t(MLP_PARAMS)) return;
    uint32_t k, o, i;
    param_position(idx, k, o, i);
    float g = 0.0f;
    for (uint32_t c = 0; c < chunks; c++) g += partial[((size_t(k) * chunks + c) * 64 + o) * 64 + i];
    if (grad_out) grad_out[idx] = g;
    if (!apply) return;
    float mm = a.beta1 * m1[idx] + (1.0f - a.beta1) * g;
    float vv = a.beta2 * m2[idx] + (1.0f - a.beta2) * (g * g);
    m1[idx] = mm;
    m2[idx] = vv;
    const float w = master[idx] - a.lr * (mm * a.c1) / (ttm::sqrt(vv * a.c2) + a.eps);
    master[idx] = w;
    const uint16_t h = float_to_bf16(w);
    if (k < 5u) {
        wf[(64u * k + o) * MLP_ROW + i] = h;
        if (k >= 1u) wb[(64u * (k - 1u) + i) * MLP_ROW + o] = h;
    } else {
        wf[(320u + o) * MLP_ROW + i] = h;
        wb5[i * MLP_BWD5_ROW + o] = h;
    }
}

// bf16 copies: forward rows [out][in] (W_5 rows 3..31 zero), transposed
// W_1^T..W_4^T [in][out] and W_5^T [in][8].
__global__ void mlp_pack_kernel(const float* master, uint16_t* wf, uint16_t* wb, uint16_t* wb5) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t fwd = MLP_FWD_ROWS * MLP_ROW, bwd = MLP_BWD_ROWS * MLP_ROW, bwd5 = 64 * MLP_BWD5_ROW;
    if (t < fwd) {
        uint32_t row = t / MLP_ROW, col = t % MLP_ROW;
        float v = 0.0f;
        if (col < 64) {
            if (row < 320) v = master[row * 64 + col];
            else if (row - 320 < 3) v = master[20480 + (row - 320) * 64 + col];
        }
        wf[t] = float_to_bf16(v);
        return;
    }
    t -= fwd;
    if (t < bwd) {
        uint32_t row = t / MLP_ROW, col = t % MLP_ROW;  // W_{k+1}^T [in = row % 64][out = col]
        uint32_t k = 1 + row / 64, in = row % 64;
        float v = col < 64 ? master[k * 4096 + col * 64 + in] : 0.0f;
        wb[t] = float_to_bf16(v);
        return;
    }
    t -= bwd;
    if (t < bwd5) {
        uint32_t in = t / MLP_BWD5_ROW, out = t % MLP_BWD5_ROW;
        wb5[t] = float_to_bf16(out < 3 ? master[20480 + out * 64 + in] : 0.0f);
    }
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
    // Training scratch, grown on demand.
    uint32_t train_capacity = 0;
    void* train_memory = nullptr;
    float* partial = nullptr;
    float* loss_tiles = nullptr;
    uint16_t* acts = nullptr;
    uint16_t* deltas = nullptr;
    float last_loss = 0.0f;
};

namespace tt {

using tt::fail;

tt_status mlp_repack(tt_mlp* m) {
    tt_context* c = m->context;
    uint32_t total = MLP_FWD_ROWS * MLP_ROW + MLP_BWD_ROWS * MLP_ROW + 64 * MLP_BWD5_ROW;
    mlp_pack_kernel<<<blocks_of(total, 256), 256, 0, tt::api_stream(c)>>>(m->master, m->wf, m->wb, m->wb5);
    TT_HIP(c, hipGetLastError());
    return TT_OK;
}

static tt_status ensure_train(tt_mlp* m, uint32_t nstride) {
    tt_context* c = m->context;
    if (nstride <= m->train_capacity) return TT_OK;
    if (m->train_memory) {
        TT_HIP(c, hipStreamSynchronize(tt::api_stream(c)));
        TT_HIP(c, hipFree(m->train_memory));
        m->train_memory = nullptr;
        m->train_capacity = 0;
    }
    size_t chunks = nstride / MLP_CHUNK;
    size_t partial = 6 * chunks * 64 * 64 * sizeof(float);
    size_t loss = (nstride / 32) * sizeof(float);
    size_t planes = 6 * size_t(64) * nstride * sizeof(uint16_t);
    TT_HIP(c, hipMalloc(&m->train_memory, partial + loss + 2 * planes));
    char* p = static_cast<char*>(m->train_memory);
    m->partial = reinterpret_cast<float*>(p);
    m->loss_tiles = reinterpret_cast<float*>(p + partial);
    m->acts = reinterpret_cast<uint16_t*>(p + partial + loss);
    m->deltas = reinterpret_cast<uint16_t*>(p + partial + loss + planes);
    m->train_capacity = nstride;
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

tt_status mlp_infer_device(tt_mlp* m, const float4* records, const float4* scale_path, const uint32_t* count_dev,
                           uint32_t count_host, float* out_rgb, float4* radiance) {
    tt_context* c = m->context;
    mlp_infer_kernel<<<c->cu_count * 2, MLP_BLOCK, 0, tt::api_stream(c)>>>(m->wf, m->bounds, records, scale_path, count_dev,
                                                                  count_host, out_rgb, radiance);
    TT_HIP(c, hipGetLastError());
    return TT_OK;
}

// Gradients (and, with apply, one Adam step) of the loss over n device
// records/targets.  grad_out (device, MLP_PARAMS floats) is optional.
tt_status mlp_train_device(tt_mlp* m, const float4* records, const float4* targets, uint32_t n, bool apply,
                           float* grad_out) {
    tt_context* c = m->context;
    if (n == 0) return TT_OK;
    uint32_t nstride = (n + MLP_CHUNK - 1) / MLP_CHUNK * MLP_CHUNK;
    tt_status st = ensure_train(m, nstride);
    if (st != TT_OK) return st;
    float inv_norm = 1.0f / (3.0f * float(n));
    uint32_t tiles = nstride / 32;
    uint32_t grid = std::min<uint32_t>(blocks_of(tiles, MLP_BLOCK / 64), uint32_t(c->cu_count));
    mlp_train_kernel<<<grid, MLP_BLOCK, 0, tt::api_stream(c)>>>(m->wf, m->wb, m->wb5, m->bounds, records, targets, n, nstride,
                                                       inv_norm, m->acts, m->deltas, m->loss_tiles);
    TT_HIP(c, hipGetLastError());
    uint32_t chunks = nstride / MLP_CHUNK;
    mlp_wgrad_kernel<<<dim3(chunks, 6), 64, 0, tt::api_stream(c)>>>(m->acts, m->deltas, nstride, m->partial);
    TT_HIP(c, hipGetLastError());
    AdamArgs a{m->lr, m->beta1, m->beta2, m->eps, 1.0f, 1.0f};
    if (apply) {
        m->step++;
        a.c1 = float(1.0 / (1.0 - std::pow(double(m->beta1), double(m->step))));
        a.c2 = float(1.0 / (1.0 - std::pow(double(m->beta2), double(m->step))));
    }
    mlp_adam_kernel<<<blocks_of(MLP_PARAMS, 256), 256, 0, tt::api_stream(c)>>>(m->master, m->m1, m->m2, m->partial, chunks, a,
                                                                      apply ? 1 : 0, grad_out, m->wf, m->wb, m->wb5);
    TT_HIP(c, hipGetLastError());
    return TT_OK;
}

tt_status mlp_last_loss(tt_mlp* m, uint32_t n, double* out) {
    tt_context* c = m->context;
    uint32_t tiles = (n + MLP_CHUNK - 1) / MLP_CHUNK * MLP_CHUNK / 32;
    std::vector<float> v(tiles);
    TT_HIP(c, hipMemcpyAsync(v.data(), m->loss_tiles, tiles * sizeof(float), hipMemcpyDeviceToHost, tt::api_stream(c)));
    TT_HIP(c, hipStreamSynchronize(tt::api_stream(c)));
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

template <typename T>
tt_status upload_host(tt_context* c, T* dst, const T* src, size_t count) {
    TT_HIP(c, hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, tt::api_stream(c)));
    TT_HIP(c, hipStreamSynchronize(tt::api_stream(c)));
    return TT_OK;
}

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
    size_t fwd = size_t(MLP_FWD_ROWS) * MLP_ROW, bwd = size_t(MLP_BWD_ROWS) * MLP_ROW, bwd5 = 64 * MLP_BWD5_ROW;
    if (hipMalloc(&m->master, 3 * MLP_PARAMS * sizeof(float)) != hipSuccess) return bail(TT_OUT_OF_MEMORY);
    m->m1 = m->master + MLP_PARAMS;
    m->m2 = m->master + 2 * MLP_PARAMS;
    if (hipMalloc(&m->wf, (fwd + bwd + bwd5) * sizeof(uint16_t)) != hipSuccess) return bail(TT_OUT_OF_MEMORY);
    m->wb = m->wf + fwd;
    m->wb5 = m->wb + bwd;
    // He-uniform initialisation: U(-sqrt(6 / 64), sqrt(6 / 64)); Adam state 0.
    std::vector<float> host(3 * MLP_PARAMS, 0.0f);
    HostPcg rng(desc ? desc->seed : 0);
    const float bound = std::sqrt(6.0f / 64.0f);
    for (int i = 0; i < MLP_PARAMS; i++) host[i] = (2.0f * rng.uniform() - 1.0f) * bound;
    tt_status st = upload_host(c, m->master, host.data(), host.size());
    if (st != TT_OK) return bail(st);
    st = tt::mlp_repack(m);
    if (st != TT_OK) return bail(st);
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
    if (count != uint32_t(MLP_PARAMS)) return fail(c, TT_INVALID_ARGUMENT, "parameter count is 20672");
    TT_HIP(c, hipSetDevice(c->device));
    if (get) {
        TT_HIP(c, hipMemcpyAsync(get, m->master, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, tt::api_stream(c)));
        TT_HIP(c, hipStreamSynchronize(tt::api_stream(c)));
    }
    if (set) {
        tt_status st = upload_host(c, m->master, set, MLP_PARAMS);
        if (st != TT_OK) return st;
        st = tt::mlp_repack(m);
        if (st != TT_OK) return st;
        TT_HIP(c, hipStreamSynchronize(tt::api_stream(c)));
    }
    return TT_OK;
} catch (...) {
    return tt::exception_status(m && m->context ? &m->context->error : nullptr);
}

// Host-array entry points (tests): op 0 encode (out: 64 floats per record),
// 1 infer (out: 3 floats), 2 gradient (targets: 4 floats per record; out:
// MLP_PARAMS floats; loss), 3 one Adam step (targets; loss).
tt_status tt_mlp_run(tt_mlp* m, int op, const float* records, const float* targets, uint32_t n, float* out,
                     double* loss) try {
    if (!m || !records || n == 0) return fail(m ? m->context : nullptr, TT_INVALID_ARGUMENT, "null argument");
    tt_context* c = m->context;
    if ((op >= 2 && !targets) || (op <= 2 && !out) || op < 0 || op > 3)
        return fail(c, TT_INVALID_ARGUMENT, "bad tt_mlp_run arguments");
    TT_HIP(c, hipSetDevice(c->device));
    size_t rec_bytes = size_t(n) * MLP_RECORD_FLOATS * sizeof(float);
    size_t tgt_bytes = size_t(n) * 4 * sizeof(float);
    size_t out_floats = op == 0 ? size_t(n) * MLP_IN : (op == 1 ? size_t(n) * 3 : (op == 2 ? MLP_PARAMS : 0));
    char* d = nullptr;
    TT_HIP(c, hipMalloc(&d, rec_bytes + tgt_bytes + out_floats * sizeof(float) + 16));
    float4* drec = reinterpret_cast<float4*>(d);
    float4* dtgt = reinterpret_cast<float4*>(d + rec_bytes);
    float* dout = reinterpret_cast<float*>(d + rec_bytes + tgt_bytes);
    tt_status st = TT_OK;
    hipError_t e = hipMemcpyAsync(drec, records, rec_bytes, hipMemcpyHostToDevice, tt::api_stream(c));
    if (e == hipSuccess && targets) e = hipMemcpyAsync(dtgt, targets, tgt_bytes, hipMemcpyHostToDevice, tt::api_stream(c));
    if (e != hipSuccess) st = tt::hip_fail(c, e, "tt_mlp_run upload");
    if (st == TT_OK) {
        if (op == 0) {
            mlp_encode_kernel<<<blocks_of(n, 64), 64, 0, tt::api_stream(c)>>>(m->bounds, drec, n, dout);
            e = hipGetLastError();
            if (e != hipSuccess) st = tt::hip_fail(c, e, "encode");
        } else if (op == 1) {
            st = tt::mlp_infer_device(m, drec, nullptr, nullptr, n, dout, nullptr);
        } else {
            st = tt::mlp_train_device(m, drec, dtgt, n, op == 3, op == 2 ? dout : nullptr);
        }
    }
    if (st == TT_OK && out_floats) {
        e = hipMemcpyAsync(out, dout, out_floats * sizeof(float), hipMemcpyDeviceToHost, tt::api_stream(c));
        if (e != hipSuccess) st = tt::hip_fail(c, e, "tt_mlp_run download");
    }
    if (st == TT_OK && op >= 2 && loss) st = tt::mlp_last_loss(m, n, loss);
    if (st == TT_OK) {
        e = hipStreamSynchronize(tt::api_stream(c));
        if (e != hipSuccess) st = tt::hip_fail(c, e, "tt_mlp_run");
    }
    (void)hipFree(d);
    return st;
} catch (...) {
    return tt::exception_status(m && m->context ? &m->context->error : nullptr);
}

}  // extern "C"
