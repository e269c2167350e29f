"""The oracle's BSDF against closed forms and invariants.

Pinned here: dielectric Fresnel against the textbook (pbrt) formula and its
normal-incidence value ((n-1)/(n+1))^2; GGX normalisation
(integral of D(m) cos(theta_m) over the hemisphere = 1, isotropic and
anisotropic); the Lambert limit of the energy-compensated Oren-Nayar lobe
(sigma = 0 gives rho/pi + 1e-7 rho^2/pi, diffuse_brdf.glsl:26-31) with
density cos/pi; uniformity of the Heitz triangle map and of cosine sampling;
sample/evaluate consistency of bsdf_sample; white-furnace energy bounds;
Oren-Nayar reciprocity.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import binding
from tinytrace_amd import Material


def fresnel_textbook(c, n):
    s2t = (1 - c * c) / (n * n)
    if s2t >= 1:
        return 1.0
    ct = math.sqrt(1 - s2t)
    rpar = (n * c - ct) / (n * c + ct)
    rper = (c - n * ct) / (c + n * ct)
    return 0.5 * (rpar * rpar + rper * rper)


def test_fresnel_dielectric():
    L = binding.lib()
    rng = np.random.default_rng(0)
    for _ in range(2000):
        c, n = float(np.float32(rng.uniform(0, 1))), float(np.float32(rng.uniform(0.5, 2.5)))
        # float32 evaluation; cancellation near total internal reflection
        assert abs(L.oracle_fresnel_dielectric(c, n) - fresnel_textbook(c, n)) < 2e-5
    for n in (1.2, 1.5, 2.0):
        assert abs(L.oracle_fresnel_dielectric(1.0, n) - ((n - 1) / (n + 1)) ** 2) < 1e-6


@pytest.mark.parametrize("ax,ay", [(0.1, 0.1), (0.4, 0.4), (0.8, 0.8), (0.2, 0.6)])
def test_ggx_normalisation(ax, ay):
    L = binding.lib()
    nt, nphi = 2048, 256
    theta = (np.arange(nt) + 0.5) * (math.pi / 2) / nt
    phi = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    total = 0.0
    m = np.zeros(3, dtype=np.float32)
    for t in theta:
        st, ct = math.sin(t), math.cos(t)
        row = 0.0
        for p in phi[::8]:
            m[:] = (st * math.cos(p), st * math.sin(p), ct)
            row += L.oracle_microfacet_distribution(ax, ay, m.ctypes.data)
        total += row * 8 * ct * st
    total *= (math.pi / 2 / nt) * (2 * math.pi / nphi)
    assert abs(total - 1.0) < 2e-2


def test_lambert_limit_of_diffuse_lobe():
    rho = np.array([0.7, 0.4, 0.2])
    mc = Material(base_color=tuple(rho), specular_weight=0.0, specular_roughness=1.0).constants()
    rng = np.random.default_rng(1)
    wi = rng.standard_normal((500, 3))
    wo = rng.standard_normal((500, 3))
    wi[:, 2], wo[:, 2] = np.abs(wi[:, 2]) + 0.05, np.abs(wo[:, 2]) + 0.05
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    r = binding.bsdf_evaluate(mc, wi, wo)
    expected = rho / math.pi + 1e-7 * rho ** 2 / math.pi
    np.testing.assert_allclose(r["bsdf"], np.broadcast_to(expected, r["bsdf"].shape), rtol=2e-6)
    np.testing.assert_allclose(r["density"], wo[:, 2] / math.pi, rtol=2e-6)
    # below the surface: zero
    r2 = binding.bsdf_evaluate(mc, wi, wo * np.array([1, 1, -1]))
    assert not r2["bsdf"].any()


def test_sample_triangle_is_uniform():
    out = np.zeros(3, dtype=np.float32)
    rng = np.random.default_rng(2)
    bary = []
    for u, v in rng.random((40000, 2)):
        binding.lib().oracle_sample_triangle(float(u), float(v), out.ctypes.data)
        bary.append(out.copy())
    b = np.array(bary)
    assert np.all(b >= -1e-7) and np.allclose(b.sum(1), 1, atol=1e-6)
    np.testing.assert_allclose(b.mean(0), [1 / 3] * 3, atol=5e-3)
    # Equal-area sub-triangles (split at the centroid) receive equal counts.
    region = np.argmin(b, axis=1)
    counts = np.bincount(region, minlength=3)
    assert np.all(np.abs(counts - b.shape[0] / 3) < 4 * math.sqrt(b.shape[0] / 3))


def test_cosine_hemisphere_sampling():
    out = np.zeros(3, dtype=np.float32)
    rng = np.random.default_rng(3)
    z = []
    for u, v in rng.random((40000, 2)):
        binding.lib().oracle_cosine_hemisphere_sample(float(u), float(v), out.ctypes.data)
        assert abs(np.linalg.norm(out) - 1) < 1e-5
        z.append(out[2])
    z = np.array(z)
    # cos-weighted: E[z] = 2/3, z^2 uniform on [0, 1]
    assert abs(z.mean() - 2 / 3) < 5e-3
    counts = np.histogram(z * z, bins=20, range=(0, 1))[0]
    assert np.all(np.abs(counts - 2000) < 5 * math.sqrt(2000))


MATERIALS = {
    "diffuse": Material(base_color=(0.7, 0.5, 0.3), specular_weight=0.0, specular_roughness=1.0),
    "plastic": Material(base_color=(0.2, 0.6, 0.3), specular_roughness=0.4),
    "metal": Material(base_color=(0.9, 0.6, 0.3), base_metalness=1.0, specular_roughness=0.3),
    "coated": Material(base_color=(0.5, 0.1, 0.1), coat_weight=1.0, specular_roughness=0.6),
    "rough_diffuse": Material(base_color=(0.6, 0.6, 0.6), base_diffuse_roughness=0.7),
    "glass": Material(transmission_weight=1.0, specular_roughness=0.2),
    "fuzz": Material(fuzz_weight=0.8, fuzz_roughness=0.4),
}


CONSISTENT = ["diffuse", "plastic", "metal", "coated", "rough_diffuse"]


@pytest.mark.parametrize("name", CONSISTENT)
def test_sample_matches_evaluate(name):
    """bsdf_sample's value and density equal bsdf_evaluate's at the sampled
    direction (same lobes, same generator state).  Holds for the reference's
    diffuse, specular-reflection, metal and coat lobes; its transmission and
    fuzz sampling routines disagree with their evaluate routines (see the two
    tests below), which the oracle reproduces."""
    mc = MATERIALS[name].constants()
    rng = np.random.default_rng(4)
    n = 4000
    wi = rng.standard_normal((n, 3))
    wi[:, 2] = np.abs(wi[:, 2]) + 0.1
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    out = binding.bsdf_sample(mc, wi, rng.integers(1, 2**32, n, dtype=np.uint64).astype(np.uint32))
    ok = out[:, 7] == 1
    assert ok.mean() > 0.5
    value, density = out[ok, 3:6], out[ok, 6]
    ev, ed = out[ok, 9:12], out[ok, 12]
    good = (density > 1e-3) & np.all(np.isfinite(value), axis=1)
    rel_v = np.abs(value[good] - ev[good]) / np.maximum(np.abs(ev[good]), 1e-4)
    rel_d = np.abs(density[good] - ed[good]) / np.maximum(ed[good], 1e-4)
    # The sampled lobe's value comes from the sampling routine (its own
    # half vector); agreement within 1e-3 for nearly all samples.
    assert np.mean(rel_v.max(axis=1) < 1e-3) > 0.99
    assert np.mean(rel_d < 1e-3) > 0.99


@pytest.mark.parametrize("name", sorted(MATERIALS))
def test_white_furnace_energy(name):
    """E[f cos / pdf] over sampled directions (the throughput factor) stays
    at most ~1 for every material at normal-ish incidence."""
    mc = MATERIALS[name].constants()
    rng = np.random.default_rng(6)
    n = 20000
    wi = np.tile([0.2, 0.1, 0.97], (n, 1)).astype(np.float32)
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    out = binding.bsdf_sample(mc, wi, rng.integers(1, 2**32, n, dtype=np.uint64).astype(np.uint32))
    ok = out[:, 7] == 1
    out = out.astype(np.float64)
    w = np.where(ok[:, None], out[:, 3:6] * np.abs(out[:, 2:3]) / np.maximum(out[:, 6:7], 1e-6), 0.0)
    mean = w.mean(0)
    assert np.all(mean < 1.05), mean
    if name == "diffuse":
        np.testing.assert_allclose(mean, [0.7, 0.5, 0.3], rtol=1e-5)


def test_oren_nayar_reciprocity():
    mc = MATERIALS["rough_diffuse"].constants()
    rng = np.random.default_rng(7)
    wi = rng.standard_normal((1000, 3))
    wo = rng.standard_normal((1000, 3))
    wi[:, 2], wo[:, 2] = np.abs(wi[:, 2]) + 0.05, np.abs(wo[:, 2]) + 0.05
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    a = binding.bsdf_evaluate(mc, wi, wo, precision=64)["bsdf"]
    b = binding.bsdf_evaluate(mc, wo, wi, precision=64)["bsdf"]
    # specular weight 1 here: the specular lobe's albedo estimate depends on
    # wi, so compare the diffuse-only material for strict reciprocity.
    mc2 = Material(base_color=(0.6, 0.6, 0.6), base_diffuse_roughness=0.7, specular_weight=0.0).constants()
    a = binding.bsdf_evaluate(mc2, wi, wo, precision=64)["bsdf"]
    b = binding.bsdf_evaluate(mc2, wo, wi, precision=64)["bsdf"]
    np.testing.assert_allclose(a, b, rtol=1e-5)


def test_reference_transmission_sample_stays_above_the_surface():
    """specular_btdf_sample (specular_btdf.glsl:95-98) negates the refracted
    direction, so its 'transmitted' samples leave on the incident side, where
    specular_btdf_evaluate returns zero (specular_btdf.glsl:29-32)."""
    mc = Material(transmission_weight=1.0, specular_roughness=0.2).constants()
    rng = np.random.default_rng(8)
    wi = rng.standard_normal((2000, 3))
    wi[:, 2] = np.abs(wi[:, 2]) + 0.1
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    out = binding.bsdf_sample(mc, wi, rng.integers(1, 2**32, 2000, dtype=np.uint64).astype(np.uint32))
    btdf = (out[:, 7] == 1) & (out[:, 8] == 4)
    assert btdf.sum() > 1000
    assert np.all(out[btdf, 2] >= 0)


def test_reference_fuzz_sample_density_scales_with_ltc_norm():
    """fuzz_brdf_sample evaluates its density at the LTC-space direction it
    drew, while fuzz_brdf_evaluate maps wo back to a direction scaled by the
    norm of the transformed sample (fuzz_brdf.glsl:44-69), so the two
    densities differ by a power of that norm."""
    mc = Material(fuzz_weight=1.0, fuzz_roughness=0.4, base_color=(0.0, 0.0, 0.0), specular_weight=0.0).constants()
    rng = np.random.default_rng(9)
    wi = np.tile([0.3, 0.0, 0.954], (2000, 1))
    out = binding.bsdf_sample(mc, wi, rng.integers(1, 2**32, 2000, dtype=np.uint64).astype(np.uint32))
    fuzz = (out[:, 7] == 1) & (out[:, 8] == 0)
    assert fuzz.sum() > 1000
    ratio = out[fuzz, 6] / out[fuzz, 12]
    assert np.mean(np.abs(ratio - 1) > 1e-3) > 0.5
