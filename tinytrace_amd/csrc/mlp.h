// Radiance-cache MLP: layout, feature encoding and the device functions the
// integrator and the MLP kernels share.
//
// Network (no biases; feature 60 is a constant 1):
//   a0 = encode(record)              64 features
//   a_l = relu(W_{l-1} a_{l-1})      l = 1..5, W_0..W_4: 64 x 64
//   L = softplus(W_5 a_5)            W_5: 3 x 64, L = radiance (RGB)
// Parameters: 5 * 64 * 64 + 3 * 64 = 20672, stored fp32 (master, Adam state)
// in layer order, each matrix row-major [out][in]; the kernels read bf16
// copies and accumulate in fp32 with v_mfma_f32_32x32x8bf16_1k.
#pragma once

#include <cstdint>

#include "tt_math.h"

namespace tt {

constexpr int MLP_IN = 64;
constexpr int MLP_WIDTH = 64;
constexpr int MLP_OUT = 3;
constexpr int MLP_LAYERS = 6;                       // W_0 .. W_5
constexpr int MLP_PARAMS = 5 * 64 * 64 + 3 * 64;     // 20672
constexpr int MLP_RECORD_FLOATS = 16;                // raw features per query
// bf16 forward weights in LDS/global: rows of 68 halves (64 + 4 pad, keeps
// the 32 rows a wave reads on distinct LDS banks); W_5 padded to 32 rows.
constexpr int MLP_ROW = 68;
constexpr int MLP_FWD_ROWS = 5 * 64 + 32;            // 352
// bf16 transposed weights for the backward pass, W_l^T [in][out] for
// l = 1..4 (64 rows of 68) and W_5^T [in][out padded to 8] (64 rows of 8).
constexpr int MLP_BWD_ROWS = 4 * 64;                  // W_1^T .. W_4^T
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

// Encoded feature f of a record (the same arithmetic on host and device):
//   0..35  position u = (p - lo) / extent clamped to [0, 1]: for axis c and
//          octave i = 0..5: sin(2^i pi u_c), cos(2^i pi u_c) at 12c + 2i + {0,1}
//   36..38 direction, 39..41 normal, 42..44 base colour
//   45 roughness, 46 metalness, 47 bounce / 8
//   48..59 direction: for axis c: sin(pi d), cos(pi d), sin(2 pi d),
//          cos(2 pi d) at 48 + 4c + {0..3}
//   60 constant 1, 61..63 zero
struct MlpFeatures {
    float f[MLP_IN];
};

#if defined(__HIPCC__)
#define TT_MLP_HD __host__ __device__ inline
#else
#define TT_MLP_HD inline
#endif

TT_MLP_HD void mlp_encode(const MlpRecord& r, const MlpBounds& b, MlpFeatures& out) {
    const float pi = 3.14159265358979323846f;
    for (int c = 0; c < 3; c++) {
        float u = (r.v[c] - b.lo[c]) * b.inv_extent[c];
        u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
        float scale = pi;
        for (int i = 0; i < 6; i++) {
            float x = scale * u;
            out.f[12 * c + 2 * i] = ttm::sin(x);
            out.f[12 * c + 2 * i + 1] = ttm::cos(x);
            scale = scale * 2.0f;
        }
    }
    for (int k = 0; k < 3; k++) {
        out.f[36 + k] = r.v[3 + k];
        out.f[39 + k] = r.v[6 + k];
        out.f[42 + k] = r.v[9 + k];
    }
    out.f[45] = r.v[12];
    out.f[46] = r.v[13];
    out.f[47] = r.v[14] * 0.125f;
    for (int c = 0; c < 3; c++) {
        float d = r.v[3 + c];
        out.f[48 + 4 * c + 0] = ttm::sin(pi * d);
        out.f[48 + 4 * c + 1] = ttm::cos(pi * d);
        out.f[48 + 4 * c + 2] = ttm::sin(2.0f * pi * d);
        out.f[48 + 4 * c + 3] = ttm::cos(2.0f * pi * d);
    }
    out.f[60] = 1.0f;
    out.f[61] = 0.0f;
    out.f[62] = 0.0f;
    out.f[63] = 0.0f;
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
