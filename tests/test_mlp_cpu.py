"""The radiance-cache MLP oracle (oracle/mlp_ref.py): its analytic gradient
against central finite differences in float64, bfloat16 rounding, and the
encoding's closed-form properties."""
import numpy as np

from oracle import mlp_ref


def _problem(n=24, seed=0):
    rng = np.random.default_rng(seed)
    params = rng.uniform(-1, 1, mlp_ref.PARAMS) * np.sqrt(6 / 64)
    records = rng.uniform(-1, 1, (n, 16))
    records[:, 12:15] = rng.uniform(0, 1, (n, 3))
    feats = mlp_ref.encode(records, (-1, -1, -1), (1, 1, 1))
    targets = rng.uniform(0, 2, (n, 3))
    return params, feats, targets


def test_gradient_matches_finite_differences():
    params, feats, targets = _problem()
    grad, _ = mlp_ref.gradient(params, feats, targets)
    rng = np.random.default_rng(1)
    # Parameters of every layer, including the 96-input first and the output layer.
    idx = np.concatenate([rng.integers(a, b, 6) for a, b in mlp_ref.layer_slices()])
    fd = mlp_ref.finite_difference(params, feats, targets, idx)
    np.testing.assert_allclose(grad[idx], fd, rtol=2e-5, atol=1e-9)


def test_bf16_points_change_the_gradient_only_slightly():
    params, feats, targets = _problem(n=64)
    g64, l64 = mlp_ref.gradient(params, feats, targets)
    g16, l16 = mlp_ref.gradient(params, feats, targets, bf16_points=True)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 0.05
    assert abs(l16 - l64) / l64 < 0.05


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 3.0e38, np.inf], dtype=np.float32)
    # 3e38 = 1.7632 * 2^127 -> 1 + 98/128 in bf16's 7 stored mantissa bits.
    expected = [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, (1 + 98 / 128) * 2.0 ** 127, np.inf]
    np.testing.assert_allclose(mlp_ref.bf16(x), expected, rtol=1e-7)


def test_encoding_properties():
    rng = np.random.default_rng(2)
    rec = rng.uniform(-3, 3, (100, 16)).astype(np.float32)
    f = mlp_ref.encode(rec, (-2, -2, -2), (2, 2, 2))
    assert f.shape == (100, 96)
    # 12 frequencies of each position axis: unit (sin, cos) pairs, and each
    # frequency the double angle of the one before it.
    for c in range(3):
        for i in range(12):
            s, co = f[:, 24 * c + 2 * i], f[:, 24 * c + 2 * i + 1]
            np.testing.assert_allclose(s * s + co * co, 1.0, atol=1e-12)
            if i:
                ps, pc = f[:, 24 * c + 2 * i - 2], f[:, 24 * c + 2 * i - 1]
                np.testing.assert_allclose(s, 2 * ps * pc, atol=1e-9)
                np.testing.assert_allclose(co, pc * pc - ps * ps, atol=1e-9)
    u = np.clip((rec[:, 0].astype(np.float64) + 2) / 4, 0, 1)
    np.testing.assert_allclose(f[:, 0], np.sin(np.pi * u), atol=1e-6)
    np.testing.assert_allclose(f[:, 22], np.sin(np.pi * 2048 * u), atol=2e-3)
    np.testing.assert_array_equal(f[:, 72:75], rec[:, 6:9])
    np.testing.assert_array_equal(f[:, 75:78], rec[:, 9:12])
    np.testing.assert_array_equal(f[:, 78:80], rec[:, 12:14])
    assert np.all(f[:, 81] == 1.0)
    # The bounce one-hot: 1, 2, >= 3.
    b = np.array([0, 1, 2, 3, 7], dtype=np.float32)
    rb = np.zeros((5, 16), dtype=np.float32)
    rb[:, 14] = b
    fb = mlp_ref.encode(rb, (-2, -2, -2), (2, 2, 2))
    np.testing.assert_array_equal(fb[:, [80, 94, 95]], [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 1]])
    # Direction: octaves 0 and 1 of each axis.
    for c in range(3):
        d = rec[:, 3 + c].astype(np.float64)
        np.testing.assert_allclose(f[:, 82 + 4 * c], np.sin(np.pi * d), atol=1e-12)
        np.testing.assert_allclose(f[:, 85 + 4 * c], np.cos(2 * np.pi * d), atol=1e-12)


def test_encoding_uses_the_kernels_float32_coordinate():
    """u is (p - lo) * (1 / extent) in float32: at 2^11 pi u a float64 u
    would differ by up to ~1e-3 from what the kernels encode."""
    rng = np.random.default_rng(3)
    rec = rng.uniform(-3, 3, (2000, 16)).astype(np.float32)
    lo, hi = (-2.3, -2.3, -2.3), (1.7, 1.7, 1.7)
    u32 = mlp_ref.unit_coordinates(rec, lo, hi)[:, 0]
    f = mlp_ref.encode(rec, lo, hi)
    np.testing.assert_array_equal(f[:, 22], np.sin(np.pi * 2048 * u32))
    lo32, ext32 = np.float32(-2.3), np.float32(1.7) - np.float32(-2.3)
    u_ref = np.clip((rec[:, 0] - lo32) * (np.float32(1) / ext32), 0, 1).astype(np.float64)
    np.testing.assert_array_equal(u32, u_ref)


def test_adam_first_step_is_signed_learning_rate():
    """With zero state, the first bias-corrected Adam step moves every
    parameter by lr * g / (|g| + eps) ~ lr * sign(g)."""
    p = np.zeros(10, dtype=np.float32)
    g = np.linspace(-1, 1, 10).astype(np.float32)
    g[4] = 0.25
    q, _, _ = mlp_ref.adam(p, g, np.zeros(10), np.zeros(10), 1)
    np.testing.assert_allclose(q, -1e-3 * np.sign(g), rtol=1e-4)
