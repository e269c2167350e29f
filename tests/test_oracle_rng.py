"""Generators of the oracle (src/includes/random.glsl:4-50 and PCG).

The reference generator is pinned by integer known-answer vectors computed
from the text of random.glsl:4-38 in Python integer arithmetic (jenkins_hash,
seeding from pixel and frame, Marsaglia xorshift32 (13, 17, 5), the
0x3f800000 | v >> 9 float).  The PCG sampler is pcg32 (PCG-XSH-RR 64/32),
pinned by the published output of the PCG reference implementation's demo;
its per-path seeding goes through SplitMix64, pinned by splitmix64.c's
published sequence.
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


# pcg32 demo of the PCG reference implementation (pcg-c-basic, pcg32-demo):
# pcg32_srandom_r(&rng, 42u, 54u), then six pcg32_random_r draws.
PCG32_42_54 = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
# splitmix64.c (Vigna) seeded with 0: the first three next() outputs, i.e.
# the finaliser at k * 0x9e3779b97f4a7c15 for k = 0, 1, 2.
SPLITMIX64_0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_pcg32_known_answers():
    assert [int(v) for v in binding.pcg32_draws(42, 54, 6)] == PCG32_42_54


def test_splitmix64_known_answers():
    for k, expected in enumerate(SPLITMIX64_0):
        assert binding.lib().oracle_splitmix64((k * 0x9E3779B97F4A7C15) & (2**64 - 1)) == expected


def test_path_generator_is_pcg32_on_the_default_stream():
    """kind 1 draws = pcg32 seeded with initstate splitmix64(((frame << 32) |
    index) ^ splitmix64(seed)) on the default stream (initseq such that
    2 * initseq + 1 = 0xda3e39cb94b95bdb)."""
    index, frame, seed = 3 + 7 * 512, 5, 1
    initstate = binding.lib().oracle_splitmix64(((frame << 32) | index) ^ binding.lib().oracle_splitmix64(seed))
    _, draws = binding.rng_draws(1, 3, 7, 512, frame, seed, 16)
    expected = binding.pcg32_draws(initstate, 0xDA3E39CB94B95BDB >> 1, 16)
    np.testing.assert_array_equal(draws, expected)


@pytest.mark.parametrize("kind", [0, 1])
def test_generator_uniformity(kind):
    _, draws = binding.rng_draws(kind, 3, 7, 512, 5, 1, 200000)
    u = np.array([binding.lib().oracle_uint_to_unit_float(int(d)) for d in draws[:20000]])
    assert u.min() >= 0.0 and u.max() < 1.0
    counts = np.histogram((draws >> 9).astype(np.float64) / 2**23, bins=64, range=(0, 1))[0]
    expected = draws.size / 64
    chi2 = ((counts - expected) ** 2 / expected).sum()
    assert chi2 < 120  # 63 dof, p ~ 1e-5
