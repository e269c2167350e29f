// Radiance-cache MLP: layout, feature encoding and the device functions the
// integrator and the MLP kernels share.
//
// Network (no biases; feature 81 is a constant 1):
//   a0 = encode(record)              96 features
//   a1 = relu(W_0 a0)                W_0: 64 x 96
//   a_l = relu(W_{l-1} a_{l-1})      l = 2..5, W_1..W_4: 64 x 64
//   L = softplus(W_5 a_5)            W_5: 3 x 64, L = radiance (RGB)
// Parameters: 64 * 96 + 4 * 64 * 64 + 3 * 64 = 22720, stored fp32 (master,
// Adam state) in layer order, each matrix row-major [out][in]; the kernels
// read bf16 copies and accumulate in fp32 on the matrix cores.
#pragma once

#include <cstdint>

#include "tt_math.h"

namespace tt {

constexpr int MLP_IN = 96;
constexpr int MLP_WIDTH = 64;
constexpr int MLP_OUT = 3;
constexpr int MLP_LAYERS = 6;                           // W_0 .. W_5
constexpr int MLP_W0 = 64 * MLP_IN;                     // 6144
constexpr int MLP_W5_OFFSET = MLP_W0 + 4 * 64 * 64;     // 22528
constexpr int MLP_PARAMS = MLP_W5_OFFSET + 3 * 64;      // 22720
constexpr int MLP_RECORD_FLOATS = 16;                   // raw features per query
// First parameter of W_k.
constexpr int mlp_layer_offset(int k) { return k == 0 ? 0 : MLP_W0 + 4096 * (k - 1); }

// bf16 forward weights in LDS/global: W_0 as 64 rows of MLP_ROW0 halves
// (96 + 4 pad), then W_1..W_4 and W_5 (padded to 32 rows) as rows of MLP_ROW
// halves (64 + 4 pad).  The pads keep the 32 rows a wave reads on distinct
// LDS banks.
constexpr int MLP_ROW0 = MLP_IN + 4;                     // 100
constexpr int MLP_ROW = 68;
constexpr int MLP_FWD_W1 = 64 * MLP_ROW0;                // first half of W_1
constexpr int MLP_FWD_W5 = MLP_FWD_W1 + 4 * 64 * MLP_ROW;  // first half of W_5
constexpr int MLP_FWD_HALVES = MLP_FWD_W5 + 32 * MLP_ROW;  // 25984
// bf16 transposed weights for the backward pass, W_l^T [in][out] for
// l = 1..4 (64 rows of MLP_ROW) and W_5^T [in][out padded to 8] (64 rows of 8).
constexpr int MLP_BWD_ROWS = 4 * 64;                     // W_1^T .. W_4^T
constexpr int MLP_BWD5_ROW = 8;

// Raw record (16 floats): position (3), direction toward the previous vertex
// (3), shading normal (3), base colour (3), specular roughness, metalness,
// bounce index, unused.
struct MlpRecord {
    float v[MLP_RECORD_FLOATS];
};

// Scene bounds for position normalisation.
struct MlpBounds {
    float lo[3];
    float inv_extent[3];
};

// Encoded features, as 48 pairs (2j, 2j + 1):
//   pairs 0..35 (features 0..71): position u = (p - lo) * inv_extent clamped
//          to [0, 1]; pair 12c + i (axis c, frequency i = 0..11) is
//          (sin(2^i pi u_c), cos(2^i pi u_c))
//   72..74 shading normal, 75..77 base colour, 78 roughness, 79 metalness,
//   80 [bounce == 1], 81 constant 1
//   pairs 41..46 (features 82..93): direction d; pair 41 + 2c + o (axis c,
//          octave o = 0, 1) is (sin(2^o pi d_c), cos(2^o pi d_c))
//   94 [bounce == 2], 95 [bounce >= 3]
// The bounce enters one-hot: a vertex's outgoing radiance depends on the
// bounces left after it, and training samples come from several bounces.
struct MlpFeatures {
    float f[MLP_IN];
};

#if defined(__HIPCC__)
#define TT_MLP_HD __host__ __device__ inline
#else
#define TT_MLP_HD inline
#endif

// Normalised coordinate of axis c (float arithmetic: subtract, multiply).
TT_MLP_HD float mlp_unit(const MlpRecord& r, const MlpBounds& b, int c) {
    float u = (r.v[c] - b.lo[c]) * b.inv_extent[c];
    return u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
}

TT_MLP_HD float mlp_pick3(int c, float a, float b, float d) { return c == 0 ? a : (c == 1 ? b : d); }

// 1 if the record's bounce index is b, else 0.
TT_MLP_HD float mlp_bounce_is(const MlpRecord& r, int b) { return r.v[14] == float(b) ? 1.0f : 0.0f; }

// Features 2j and 2j + 1.  The frequency 2^i multiplies exactly, and
// sincospi reduces exactly, so the highest frequency carries no more
// argument error than the lowest.  (Selects rather than indexing: on the
// device j differs between the two halves of a wave.)
TT_MLP_HD void mlp_encode_pair(const MlpRecord& r, const MlpBounds& b, int j, float& x0, float& x1) {
    if (j < 36) {
        int c = j / 12, i = j % 12;
        float u = mlp_pick3(c, mlp_unit(r, b, 0), mlp_unit(r, b, 1), mlp_unit(r, b, 2));
        ttm::sincospi(u * ttm::exp2i(i), x0, x1);
    } else if (j < 41) {
        // normal (x, y | z), base colour (r | g, b), roughness, metalness,
        // [bounce == 1], constant 1
        int k = j - 36;
        x0 = k == 0 ? r.v[6] : (k == 1 ? r.v[8] : (k == 2 ? r.v[10] : (k == 3 ? r.v[12] : mlp_bounce_is(r, 1))));
        x1 = k == 0 ? r.v[7] : (k == 1 ? r.v[9] : (k == 2 ? r.v[11] : (k == 3 ? r.v[13] : 1.0f)));
    } else if (j < 47) {
        int c = (j - 41) / 2, o = (j - 41) % 2;
        float d = mlp_pick3(c, r.v[3], r.v[4], r.v[5]);
        ttm::sincospi(o ? d * 2.0f : d, x0, x1);
    } else {
        x0 = mlp_bounce_is(r, 2);
        x1 = r.v[14] >= 3.0f ? 1.0f : 0.0f;
    }
}

TT_MLP_HD void mlp_encode(const MlpRecord& r, const MlpBounds& b, MlpFeatures& out) {
#pragma unroll
    for (int j = 0; j < MLP_IN / 2; j++) mlp_encode_pair(r, b, j, out.f[2 * j], out.f[2 * j + 1]);
}

// fp32 -> bf16, round to nearest even (NaN kept quiet).
TT_MLP_HD uint16_t float_to_bf16(float x) {
    uint32_t u = ttm::float_bits(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

TT_MLP_HD float bf16_to_float(uint16_t h) { return ttm::bits_float(uint32_t(h) << 16); }

// softplus(z) = log(1 + exp(z)), exact branch for large z.
TT_MLP_HD float mlp_softplus(float z) { return z > 20.0f ? z : ttm::log(1.0f + ttm::exp(z)); }
TT_MLP_HD float mlp_sigmoid(float z) { return 1.0f / (1.0f + ttm::exp(-z)); }

// Loss of one sample channel, relative L2 with the denominator held
// constant (Mueller et al. 2021): (y - t)^2 / (y^2 + 0.01).
constexpr float MLP_LOSS_EPS = 0.01f;

}  // namespace tt
