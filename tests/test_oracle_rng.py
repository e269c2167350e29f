"""Generators of the oracle (src/includes/random.glsl:4-50 and PCG).

The reference generator is pinned by integer known-answer vectors computed
from the text of random.glsl:4-38 in Python integer arithmetic (jenkins_hash,
seeding from pixel and frame, Marsaglia xorshift32 (13, 17, 5), the
0x3f800000 | v >> 9 float).  PCG is pinned by the invertibility of its
output permutation, the Hull-Dobell full-period conditions of its LCG, and
uniformity.
"""
import numpy as np
import pytest

from oracle import binding

JENKINS = {0: 0x00000000, 1: 0x124EA49D, 2: 0x249DC93B, 12345: 0x2BC3C9E3, 0xFFFFFFFF: 0xAE65A494}


@pytest.mark.parametrize("x,expected", sorted(JENKINS.items()))
def test_jenkins_hash_known_answers(x, expected):
    assert binding.lib().oracle_jenkins_hash(x) == expected


def test_pixel_seed_and_xorshift_states():
    # W = 512, frame 0, pixel (1, 0): seed 0x124ea49d, then three states.
    seed, draws = binding.rng_draws(0, 1, 0, 512, 0, 0, 3)
    assert seed == 0x124EA49D
    assert [hex(d) for d in draws] == ["0x1d719993", "0xe63e38f2", "0x52d6422"]
    floats = [binding.lib().oracle_uint_to_unit_float(int(d)) for d in draws]
    np.testing.assert_array_equal(np.float32(floats),
                                  np.float32([0.11501455307006836, 0.8993868827819824, 0.02022385597229004]))


def test_xorshift_fixed_point_at_pixel_zero_frame_zero():
    # jenkins_hash(0 ^ jenkins_hash(0)) = 0 and xorshift32 maps 0 to 0.
    seed, draws = binding.rng_draws(0, 0, 0, 512, 0, 0, 16)
    assert seed == 0 and not draws.any()


def test_frame_one_pixel_zero():
    seed, draws = binding.rng_draws(0, 0, 0, 512, 1, 0, 3)
    assert seed == 0xC0738807
    assert [hex(d) for d in draws] == ["0x9f15277e", "0x44a5aae3", "0xcecff1ff"]


def _pcg_output_inverse(word_out):
    # Invert out = (w >> 22) ^ w, then w = (x ^ (s >> shift)) * M with
    # M = 277803737, x = s, shift = (s >> 28) + 4.
    w = word_out
    w = w ^ (w >> 22)
    w &= 0xFFFFFFFF
    m_inv = pow(277803737, -1, 1 << 32)
    x = (w * m_inv) & 0xFFFFFFFF
    # x = s ^ (s >> k) with k = (s >> 28) + 4 >= 4: the top 4 bits of s are
    # unchanged, so k is known; then undo the xorshift.
    k = (x >> 28) + 4
    s = x
    for _ in range(32 // k + 1):
        s = x ^ (s >> k)
    return s & 0xFFFFFFFF


def test_pcg_output_is_a_bijection():
    rng = np.random.default_rng(0)
    for s in rng.integers(0, 2**32, 2000, dtype=np.uint64):
        out = binding.lib().oracle_pcg_hash((int(s) - 2891336453) * pow(747796405, -1, 1 << 32) & 0xFFFFFFFF)
        assert _pcg_output_inverse(out) == int(s)


def test_pcg_lcg_has_full_period_constants():
    # Hull-Dobell for modulus 2^32: odd increment, multiplier = 1 mod 4.
    assert 2891336453 % 2 == 1 and 747796405 % 4 == 1


@pytest.mark.parametrize("kind", [0, 1])
def test_generator_uniformity(kind):
    _, draws = binding.rng_draws(kind, 3, 7, 512, 5, 1, 200000)
    u = np.array([binding.lib().oracle_uint_to_unit_float(int(d)) for d in draws[:20000]])
    assert u.min() >= 0.0 and u.max() < 1.0
    counts = np.histogram((draws >> 9).astype(np.float64) / 2**23, bins=64, range=(0, 1))[0]
    expected = draws.size / 64
    chi2 = ((counts - expected) ** 2 / expected).sum()
    assert chi2 < 120  # 63 dof, p ~ 1e-5
