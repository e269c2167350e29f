"""Pins of the oracle's BSDF functions that sample/evaluate consistency
cannot see (both sides share the Fresnel terms and lobe weights):

* fresnel_metal (F82-tint, metal_brdf.glsl:9-15): F(1) = F0, F(0) = 1,
  F(1/7) = tint * Schlick(F0, 1/7) (the defining property of Kutz 2021),
  tint = 1 gives Schlick;
* specular_ior (Eq. 60), specular_ior_ratio (Eq. 26),
  specular_fresnel_with_coat (Eq. 75) (specular_brdf.glsl:8-28): their
  special cases, F0(ratio) = weight * F0(ior), Snell's law through the coat;
* average_dielectric_fresnel (Eqs. 102-103, fresnel.glsl:31-39): against
  2 * int_0^1 F(mu) mu dmu by quadrature, inside and outside;
* oren_nayar and energy_compensated_oren_nayar (diffuse_brdf.glsl:9-32):
  the closed-form albedo against quadrature of the single-scattering term,
  the average albedo A + (2/3 - 64/(45 pi)) B, a white furnace (albedo 1 at
  rho = 1) and the geometric-series average albedo rho E / (1 - rho (1 - E));
* the fuzz LTC (fuzz_brdf.glsl:18-42): its density integrates to 1 and
  the lobe's directional albedo is fuzz_albedo; its sample (fuzz_brdf.glsl:
  55-69) draws wo from that density and reports density and value at the
  LTC-space direction, evaluate's times |M^-1 wo|^3 and |M^-1 wo|^4;
* specular_btdf_evaluate (specular_btdf.glsl:8-54) at wo = wi refracted
  through the macro surface, where its Jacobian is Walter 2007's
  d omega_h / d omega_i (finite differences) and density and value reduce to
  textbook GGX, Smith and Fresnel factors, on both sides of the surface;
* the microfacet lobes at roughness 0: their albedo estimates reduce to
  the Fresnel factor (F82-tint, dielectric reflectance, transmittance);
* lobe_weights / lobe_probabilities (bsdf.glsl:33-113): OpenPBR's layering
  in the cases where it reduces to a few terms (Eqs. 65-71, 81, 92);
* the lobes' single-sample albedo estimators: their mean is the quadrature
  of evaluate x cos (metal, specular, coat);
* the fitted constants (Zeltner 2022, OpenPBR Eq. 102, Oren-Nayar, F82) of
  the oracle and the device against the reference's literals
  (tests/golden/bsdf_fit_constants.json).

Each pin is a function that raises on a mismatch; test_planted_slips builds
bsdf_hooks.cpp against copies of oracle_bsdf.hpp with one plausible slip
each (a power, a sign, a dropped term) and shows the pin that covers it
fails.
"""
import json
import math
import os
import re
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import binding
from tinytrace_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# Quadrature: Gauss-Legendre in mu = cos(theta), midpoint in phi.
_X, _W = np.polynomial.legendre.leggauss(160)
MU, W_MU = 0.5 * (_X + 1.0), 0.5 * _W
N_PHI = 192
PHI = (np.arange(N_PHI) + 0.5) * 2.0 * np.pi / N_PHI
W_PHI = 2.0 * np.pi / N_PHI


def hemisphere():
    """Directions (n, 3) over the upper hemisphere and their quadrature
    weights (solid angle)."""
    mu, phi = np.meshgrid(MU, PHI, indexing="ij")
    s = np.sqrt(1.0 - mu * mu)
    dirs = np.stack([s * np.cos(phi), s * np.sin(phi), mu], axis=-1).reshape(-1, 3)
    weights = np.outer(W_MU, np.full(N_PHI, W_PHI)).reshape(-1)
    return dirs, weights


def direction(mu):
    return np.array([math.sqrt(1.0 - mu * mu), 0.0, mu])


def material(**kw):
    """MaterialConstants with asset::Material's defaults and `kw` on top."""
    m = _abi.MaterialConstants()
    base = dict(base_weight=1.0, base_color=(0.8, 0.8, 0.8), specular_weight=1.0, specular_color=(1, 1, 1),
                specular_roughness=0.3, specular_ior=1.5, transmission_color=(1, 1, 1), coat_color=(1, 1, 1),
                coat_ior=1.6, coat_darkening=1.0, fuzz_color=(1, 1, 1), fuzz_roughness=0.5,
                emission_color=(1, 1, 1), geometry_normal=(0, 0, 1), geometry_opacity=1.0)
    base.update(kw)
    for k, v in base.items():
        if isinstance(v, tuple):
            getattr(m, k)[:3] = v
        else:
            setattr(m, k, v)
    return m


# ---------------------------------------------------------- textbook forms --
def schlick(f0, mu):
    return f0 + (1.0 - f0) * (1.0 - mu) ** 5


def f0_of(ior):
    return ((ior - 1.0) / (ior + 1.0)) ** 2


def fresnel_pbrt(c, n):
    """Unpolarised dielectric Fresnel reflectance (pbrt 3rd ed. 8.2.1) for
    relative IOR n, 1 under total internal reflection."""
    c = np.asarray(c, dtype=np.float64)
    s2t = (1.0 - c * c) / (n * n)
    ct = np.sqrt(np.maximum(0.0, 1.0 - s2t))
    par = (n * c - ct) / (n * c + ct)
    per = (c - n * ct) / (c + n * ct)
    return np.where(s2t >= 1.0, 1.0, 0.5 * (par * par + per * per))


def average_fresnel_quadrature(n):
    return 2.0 * float(np.sum(W_MU * fresnel_pbrt(MU, n) * MU))


# ------------------------------------------------------------------- pins --
def check_fresnel_metal(h):
    rng = np.random.default_rng(5)
    tol = 1e-6 if h.p else 2e-6
    for _ in range(20):
        f0, tint = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3)
        np.testing.assert_allclose(h.fresnel_metal(1.0, f0, tint), f0, rtol=tol, atol=tol)
        np.testing.assert_allclose(h.fresnel_metal(0.0, f0, tint), 1.0, rtol=tol)
        np.testing.assert_allclose(h.fresnel_metal(1.0 / 7.0, f0, tint), tint * schlick(f0, 1.0 / 7.0),
                                   rtol=1e-5, atol=tol)
        for mu in (0.05, 0.3, 0.6, 0.9):
            np.testing.assert_allclose(h.fresnel_metal(mu, f0, np.ones(3)), schlick(f0, mu), rtol=1e-5, atol=tol)


def check_specular_ior(h):
    for n, nc in ((1.5, 1.6), (1.33, 2.0), (2.4, 1.3)):
        assert h.specular_ior(material(specular_ior=n, coat_ior=nc, coat_weight=0.0)) == pytest.approx(n, rel=1e-6)
        assert h.specular_ior(material(specular_ior=n, coat_ior=nc, coat_weight=1.0)) == pytest.approx(n / nc, rel=1e-6)
        assert h.specular_ior(material(specular_ior=n, coat_ior=nc, coat_weight=0.5)) == \
            pytest.approx(n / (0.5 + 0.5 * nc), rel=1e-6)


def check_specular_ior_ratio(h):
    for n in (1.2, 1.5, 2.4, 0.7, 0.9):
        m = lambda w: material(specular_ior=n, specular_weight=w, coat_weight=0.0)
        assert h.specular_ior_ratio(m(1.0)) == pytest.approx(n, rel=1e-5)
        assert h.specular_ior_ratio(m(0.0)) == pytest.approx(1.0, abs=1e-7)
        for w in (0.1, 0.35, 0.8):
            r = h.specular_ior_ratio(m(w))
            assert f0_of(r) == pytest.approx(w * f0_of(n), rel=1e-4)
            assert (r - 1.0) * (n - 1.0) > 0  # the ratio stays on the ior's side of 1


def check_specular_fresnel_with_coat(h):
    for n, ior, w in ((1.5, 1.2, 0.0), (1.5, 1.2, 1.0), (1.8, 1.1, 0.4)):
        mix = n * (1 - w) + ior * w
        for c in (0.1, 0.5, 0.95):
            # coat_ior 1: the angle is unchanged.
            m = material(specular_ior=n, coat_ior=1.0, coat_weight=w)
            assert h.specular_fresnel_with_coat(m, c, ior) == pytest.approx(float(fresnel_pbrt(c, mix)), rel=1e-5)
            # Through a coat of ior 1.6: Snell's law sin_t = sin_i / 1.6.
            m = material(specular_ior=n, coat_ior=1.6, coat_weight=w)
            ct = math.sqrt(1.0 - (1.0 - c * c) / 1.6 ** 2)
            assert h.specular_fresnel_with_coat(m, c, ior) == pytest.approx(float(fresnel_pbrt(ct, mix)), rel=1e-5)


def check_average_dielectric_fresnel(h):
    # OpenPBR's fit (Eq. 102, with Eq. 103 below 1) is within 0.0015 of the
    # hemispherical average over this range.
    for n in (1.05, 1.2, 1.33, 1.5, 1.6, 2.0, 2.5, 3.0, 0.4, 0.6, 0.75, 0.9):
        assert abs(h.average_dielectric_fresnel(n) - average_fresnel_quadrature(n)) < 2e-3, n
    assert h.average_dielectric_fresnel(1.0) == 0.0


def oren_nayar_ab(s2):
    return 1.0 - 0.5 * s2 / (s2 + 0.33), 0.45 * s2 / (s2 + 0.09)


def check_oren_nayar(h):
    dirs, wts = hemisphere()
    for sigma in (0.3, 1.0, math.pi / 2):
        s2 = sigma * sigma
        a, b = oren_nayar_ab(s2)
        # Closed-form directional albedo = quadrature of the single-scattering
        # term (rho / pi)(A + B max(0, s) / max(cos_i, cos_o)), rho = 1.
        for mu in (0.1, 0.4, 0.75, 0.98):
            wi = direction(mu)
            s = dirs @ wi - dirs[:, 2] * mu
            f = (a + b * np.where(s > 0, s / np.maximum(mu, dirs[:, 2]), 0.0)) / np.pi
            want = float(np.sum(wts * f * dirs[:, 2]))
            assert h.oren_nayar(mu, s2) == pytest.approx(want, rel=2e-5), (sigma, mu)
        # Average albedo 2 int E(mu) mu dmu = A + (2/3 - 64 / (45 pi)) B.
        e = np.array([h.oren_nayar(mu, s2) for mu in MU])
        avg = 2.0 * float(np.sum(W_MU * e * MU))
        assert avg == pytest.approx(a + (2.0 / 3.0 - 64.0 / (45.0 * math.pi)) * b, rel=2e-5)


def eon_albedo(h, rho, sigma, mu, dirs, wts):
    f = h.energy_compensated_oren_nayar(np.full(3, rho), sigma, direction(mu), dirs)[:, 0]
    return float(np.sum(wts * f * dirs[:, 2]))


def check_energy_compensated_oren_nayar(h):
    dirs, wts = hemisphere()
    for sigma in (0.4, math.pi / 2):
        # White furnace: at rho = 1 nothing is lost (Eqs. 38-39).
        for mu in (0.15, 0.5, 0.9):
            assert eon_albedo(h, 1.0, sigma, mu, dirs, wts) == pytest.approx(1.0, abs=3e-4), (sigma, mu)
        # rho < 1: the cosine-weighted average of the directional albedo is
        # the sum of the scattering series rho E + rho^2 E (1 - E) + ...
        # = rho E / (1 - rho (1 - E)), E the single-scattering average.
        a, b = oren_nayar_ab(sigma * sigma)
        e_avg = a + (2.0 / 3.0 - 64.0 / (45.0 * math.pi)) * b
        for rho in (0.3, 0.7):
            mus = MU[::8]
            total = np.array([eon_albedo(h, rho, sigma, mu, dirs, wts) for mu in mus])
            avg = 2.0 * float(np.sum(W_MU[::8] * 8 * total * mus))
            assert avg == pytest.approx(rho * e_avg / (1.0 - rho * (1.0 - e_avg)), rel=5e-3), (sigma, rho)


def check_fuzz_ltc(h):
    """The LTC density integrates to 1 over the hemisphere (the transform
    keeps z, so no mass crosses the horizon), and since f cos = albedo x
    density the lobe's directional albedo is fuzz_color x fuzz_albedo."""
    dirs, wts = hemisphere()
    color = np.array([0.5, 0.7, 0.9])
    for r in (0.05, 0.3, 0.6, 1.0):
        m = material(fuzz_weight=1.0, fuzz_roughness=r, fuzz_color=tuple(color))
        for mu in (0.2, 0.6, 0.95):
            f = h.lobe_evaluate(m, "fuzz", direction(mu), dirs)
            assert float(np.sum(wts * f[:, 3])) == pytest.approx(1.0, abs=2e-3), (r, mu)
            albedo = np.sum(wts[:, None] * f[:, :3] * dirs[:, 2:3], axis=0)
            np.testing.assert_allclose(albedo, color * h.fuzz_albedo(mu, max(r, 0.01)), rtol=2e-3, atol=1e-6)


def check_smooth_limits(h):
    """At roughness 0 (alpha = 1e-4) every microfacet lobe's albedo estimate
    is its Fresnel factor at the incident angle: metal the F82-tint value,
    specular and coat the dielectric reflectance, the BTDF the dielectric
    transmittance.  The specular lobe's angle is refracted through the coat
    even at coat weight 0 (Eq. 75 as specular_brdf.glsl:23-28 applies it)."""
    states = np.random.default_rng(3).integers(1, 2 ** 32 - 1, 64, dtype=np.uint64).astype(np.uint32)
    f0, tint = np.array([0.9, 0.5, 0.2]), np.array([0.8, 0.9, 1.0])
    for mu in (0.45, 0.7, 0.95):
        wi = direction(mu)
        est = lambda lobe, **kw: h.lobe_albedo(material(**kw), lobe, wi, states).mean(axis=0)
        np.testing.assert_allclose(est("metal", base_metalness=1.0, specular_roughness=0.0, base_color=tuple(f0),
                                       specular_color=tuple(tint)), h.fresnel_metal(mu, f0, tint), rtol=5e-3)
        np.testing.assert_allclose(est("specular_brdf", specular_roughness=0.0, coat_ior=1.0)[0],
                                   fresnel_pbrt(mu, 1.5), rtol=5e-3)
        snell = math.sqrt(1.0 - (1.0 - mu * mu) / 1.6 ** 2)
        np.testing.assert_allclose(est("specular_brdf", specular_roughness=0.0, coat_ior=1.6)[0],
                                   fresnel_pbrt(snell, 1.5), rtol=5e-3)
        np.testing.assert_allclose(est("coat", coat_weight=1.0, coat_roughness=0.0, coat_ior=1.6)[0],
                                   fresnel_pbrt(mu, 1.6), rtol=5e-3)
        np.testing.assert_allclose(est("specular_btdf", transmission_weight=1.0, specular_roughness=0.0)[0],
                                   1.0 - fresnel_pbrt(mu, 1.5), rtol=5e-3)


def lobe(h, **kw):
    w, a, p = h.lobe_weights(material(**kw), direction(0.63))
    return w, a, p


def check_lobe_weights(h):
    F, C, M, S, T, D = range(6)
    tol = dict(rtol=1e-6, atol=1e-7) if h.p else dict(rtol=1e-5, atol=1e-6)
    # No fuzz, no coat: metal m, specular 1 - m, BTDF (1 - m) t, diffuse
    # (1 - m)(1 - t)(1 - E_spec).
    for m_, t in ((0.0, 0.0), (0.3, 0.0), (0.0, 0.4), (0.6, 0.25)):
        w, a, _ = lobe(h, base_metalness=m_, transmission_weight=t)
        np.testing.assert_allclose(w[F], 0.0, atol=1e-12)
        np.testing.assert_allclose(w[C], 0.0, atol=1e-12)
        np.testing.assert_allclose(w[M], m_, **tol)
        np.testing.assert_allclose(w[S], 1 - m_, **tol)
        np.testing.assert_allclose(w[T], (1 - m_) * t, **tol)
        np.testing.assert_allclose(w[D], (1 - m_) * (1 - t) * (1 - a[S]), **tol)
    # Fully metallic: no dielectric lobe at all.
    w, a, _ = lobe(h, base_metalness=1.0, transmission_weight=0.5, coat_weight=0.5)
    np.testing.assert_allclose(w[[S, T, D]], 0.0, atol=1e-12)
    np.testing.assert_allclose(a[[S, T, D]], 0.0, atol=1e-12)
    # Fuzz on an uncoated dielectric (Eq. 81): the base keeps 1 - f E_fuzz.
    for f in (0.3, 1.0):
        w, a, _ = lobe(h, fuzz_weight=f, fuzz_roughness=0.4)
        np.testing.assert_allclose(a[F], h.fuzz_albedo(0.63, 0.4), **tol)
        np.testing.assert_allclose(w[F], f, **tol)
        np.testing.assert_allclose(w[S], 1 - f * a[F], **tol)
    # Coat without darkening (Eq. 92): the base keeps coat_color (1 - E_coat).
    cc = (0.9, 0.6, 0.3)
    w, a, _ = lobe(h, coat_weight=1.0, coat_darkening=0.0, coat_color=cc)
    np.testing.assert_allclose(w[C], 1.0, **tol)
    np.testing.assert_allclose(w[S], np.array(cc) * (1 - a[C]), **tol)
    # Eq. 65 with a black base (albedo 0): darkening 1 - K.  Specular weight
    # 0 makes the base roughness 1 (Eqs. 69-70), so K is the coat's average
    # internal Fresnel (Eq. 66), the hemispherical average at ior 1 / 1.6.
    w, a, _ = lobe(h, coat_weight=1.0, coat_darkening=1.0, coat_color=cc, base_color=(0, 0, 0), specular_weight=0.0)
    k = average_fresnel_quadrature(1.0 / 1.6)
    np.testing.assert_allclose(a[D], 0.0, atol=1e-12)
    np.testing.assert_allclose(w[S], (1 - k) * np.array(cc) * (1 - a[C]), rtol=0, atol=3e-3)
    # Eq. 65 on a metal base with specular roughness 0 (base roughness 0,
    # K = F(cos wi) of the coat), partly applied (Eq. 71).
    cw, cd = 0.7, 0.6
    w, a, _ = lobe(h, base_metalness=1.0, specular_roughness=0.0, coat_weight=cw, coat_darkening=cd,
                   coat_color=cc, base_color=(0.9, 0.5, 0.2))
    kf = float(fresnel_pbrt(0.63, 1.6))
    darkening = 1 + cw * cd * ((1 - kf) / (1 - a[M] * kf) - 1)
    coated = 1 + cw * (darkening * np.array(cc) * (1 - a[C]) - 1)
    np.testing.assert_allclose(w[M], coated, rtol=1e-5, atol=1e-6)
    # Probabilities proportional to |weight x albedo|.
    for kw in (dict(base_metalness=0.4, coat_weight=0.5, fuzz_weight=0.3), dict(transmission_weight=0.6)):
        w, a, p = lobe(h, **kw)
        want = np.linalg.norm(w * a, axis=1)
        np.testing.assert_allclose(p, want / want.sum(), rtol=1e-5, atol=1e-7)
        assert p.sum() == pytest.approx(1.0, abs=1e-6)


def check_albedo_estimators(h):
    """Each lobe's single-sample albedo estimate averages to the quadrature
    of its evaluate x cos (metal, specular, coat).  The BTDF is left out: the
    reference's own sample and evaluate disagree (the sampled wo is the
    negated refraction, on wi's side, where evaluate returns 0;
    specular_btdf.glsl:28-34, 92-93), and the oracle keeps that."""
    dirs, wts = hemisphere()
    states = np.random.default_rng(9).integers(1, 2 ** 32 - 1, 40000, dtype=np.uint64).astype(np.uint32)
    cases = [("metal", dict(base_metalness=1.0, specular_roughness=0.4), dirs),
             ("specular_brdf", dict(specular_roughness=0.4), dirs),
             ("coat", dict(coat_weight=1.0, coat_roughness=0.35), dirs)]
    for name, kw, domain in cases:
        m = material(**kw)
        wi = direction(0.6)
        f = h.lobe_evaluate(m, name, wi, domain)[:, 0]
        want = float(np.sum(wts * f * np.abs(domain[:, 2])))
        est = h.lobe_albedo(m, name, wi, states)[:, 0]
        err = est.std() / math.sqrt(est.size)
        assert abs(est.mean() - want) < 4 * err + 1e-4, (name, est.mean(), want, err)


def ggx_lambda(alpha, v):
    """Smith Lambda of the isotropic GGX distribution (Heitz 2014, Eq. 72)."""
    t2 = (v[0] ** 2 + v[1] ** 2) / v[2] ** 2
    return 0.5 * (math.sqrt(1.0 + alpha * alpha * t2) - 1.0)


def half_vector_jacobian(wi, wo, eta_i, eta_o, eps=1e-5):
    """d omega_h / d omega_i of the transmission half vector
    h = normalize(-(eta_i wi + eta_o wo)) at fixed wo, by central differences
    of h over wi's spherical angles: |dh/dtheta x dh/dphi| / sin(theta)."""
    theta, phi = math.acos(wi[2]), math.atan2(wi[1], wi[0])

    def h(t, p):
        w = np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)])
        v = -(eta_i * w + eta_o * np.asarray(wo))
        return v / np.linalg.norm(v)

    dt = (h(theta + eps, phi) - h(theta - eps, phi)) / (2.0 * eps)
    dp = (h(theta, phi + eps) - h(theta, phi - eps)) / (2.0 * eps)
    return float(np.linalg.norm(np.cross(dt, dp))) / math.sin(theta)


def check_btdf_refraction(h):
    """specular_btdf_evaluate where wo is wi refracted through the macro
    surface (Snell's law, opposite azimuth), so the transmission half vector
    is the normal n.  There the reference's Jacobian
    ior^2 |cos wi| / (cos wo + ior cos wi)^2 (specular_btdf.glsl:12-13) is
    Walter et al. 2007's d omega_h / d omega_i, measured here by finite
    differences of the half-vector map, and:
      wi outside: density = G1(wi) D(n) J, the GGX visible-normal density of
        n (Heitz 2018, Eq. 3) times J;
      wi inside: density = 0 (the reference's max(0, wi.h) with h turned to
        +z, specular_btdf.glsl:16,45-46);
      both: value = (1 - F) G2 D(n) J / |cos wo| (specular_btdf.glsl:21-23 at
        h = n), F the dielectric reflectance at wi's side, D, G1, G2 the
        isotropic GGX and height-correlated Smith forms."""
    tol = 1e-7 if h.p else 1e-4
    eta = 1.5
    for r in (0.25, 0.5, 0.8):
        r = float(np.float32(r))  # MaterialConstants holds float
        alpha = r * r
        d_n = 1.0 / (math.pi * alpha * alpha)
        m = material(transmission_weight=1.0, specular_roughness=r, specular_ior=eta)
        for inside, theta in ((False, 0.2), (False, 0.7), (False, 1.2), (True, 0.15), (True, 0.5), (True, 0.7)):
            eta_i, eta_o = (eta, 1.0) if inside else (1.0, eta)
            phi = 0.4 + theta
            s_o = eta_i / eta_o * math.sin(theta)
            c_i, c_o = math.cos(theta), math.sqrt(1.0 - s_o * s_o)
            sign = -1.0 if inside else 1.0
            wi = np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), sign * c_i])
            wo = np.array([-s_o * math.cos(phi), -s_o * math.sin(phi), -sign * c_o])
            j = half_vector_jacobian(wi, wo, eta_i, eta_o)
            f = h.lobe_evaluate(m, "specular_btdf", wi, wo)[0]
            g1 = 1.0 / (1.0 + ggx_lambda(alpha, wi))
            g2 = 1.0 / (1.0 + ggx_lambda(alpha, wi) + ggx_lambda(alpha, wo))
            fresnel = float(fresnel_pbrt(c_i, eta_o / eta_i))
            want_density = 0.0 if inside else g1 * d_n * j
            assert f[3] == pytest.approx(want_density, rel=tol, abs=1e-12), (r, inside, theta)
            np.testing.assert_allclose(f[:3], (1.0 - fresnel) * g2 * d_n * j / c_o, rtol=tol, err_msg=str((r, inside, theta)))


def check_fuzz_sample(h):
    """fuzz_brdf_sample (fuzz_brdf.glsl:55-69) draws d cosine-distributed in
    the LTC's original space and returns wo = normalize(M d) in wi's LTC
    frame, M the inverse of M^-1 = [[a, 0, b], [0, a, 0], [0, 0, 1]] (a, b
    from ltc_inverse_coeffs).  So (1) the sampled wo follow the LTC density
    that evaluate returns: sample means of x, z, x^2, z^2, x z equal its
    quadrature; and (2) the reference reports density and value at the unit
    d itself, not at M^-1 wo = d / |M d|, so against evaluate at the same wo
    they carry the LTC Jacobian (a / |M^-1 wo|^2)^2 at |M^-1 wo| = 1: the
    sample's density is evaluate's times |M^-1 wo|^3 and its value
    evaluate's times |M^-1 wo|^4."""
    dirs, wts = hemisphere()
    states = np.random.default_rng(11).integers(1, 2 ** 32 - 1, 16000, dtype=np.uint64)
    tol = 1e-9 if h.p else 2e-5
    for r, mu in ((0.3, 0.5), (0.7, 0.85), (1.0, 0.3)):
        r = float(np.float32(r))  # MaterialConstants holds float
        m = material(fuzz_weight=1.0, fuzz_roughness=r, fuzz_color=(0.5, 0.7, 0.9))
        wi = direction(mu)  # azimuth 0: the LTC frame is the local frame
        a, b = h.ltc_inverse_coeffs(mu, r)
        samples = []
        for s in states:
            ok, wo, f = h.lobe_sample(m, "fuzz", wi, s)
            assert ok
            samples.append(wo)
            e = h.lobe_evaluate(m, "fuzz", wi, wo)[0]
            k = float(np.linalg.norm([a * wo[0] + b * wo[2], a * wo[1], wo[2]]))
            assert f[3] == pytest.approx(e[3] * k ** 3, rel=tol), (r, mu, wo)
            np.testing.assert_allclose(f[:3], e[:3] * k ** 4, rtol=tol)
        samples = np.array(samples)
        density = h.lobe_evaluate(m, "fuzz", wi, dirs)[:, 3]
        for g in (lambda v: v[:, 0], lambda v: v[:, 2], lambda v: v[:, 0] ** 2, lambda v: v[:, 2] ** 2,
                  lambda v: v[:, 0] * v[:, 2]):
            want = float(np.sum(wts * density * g(dirs)))
            est = g(samples)
            assert abs(est.mean() - want) < 4.5 * est.std() / math.sqrt(est.size) + 1e-3, (r, mu, est.mean(), want)


PINS = [check_fresnel_metal, check_specular_ior, check_specular_ior_ratio, check_specular_fresnel_with_coat,
        check_average_dielectric_fresnel, check_oren_nayar, check_energy_compensated_oren_nayar, check_fuzz_ltc,
        check_smooth_limits, check_lobe_weights, check_albedo_estimators, check_btdf_refraction, check_fuzz_sample]


@pytest.mark.parametrize("precision", [1, 0], ids=["double", "float"])
@pytest.mark.parametrize("pin", PINS, ids=[p.__name__[6:] for p in PINS])
def test_oracle_bsdf_pin(pin, precision):
    if pin is check_albedo_estimators and precision == 0:
        pytest.skip("Monte Carlo pin: the double instantiation covers the estimator")
    pin(binding.BsdfHooks(precision=precision))


# --------------------------------------------------------------- literals --
LITERAL = re.compile(r"(?<![\w.])(\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?f?(?![\w.])")


def function_literals(text, name):
    m = re.search(r"^[^\n;=()]*\b" + name + r"\s*\([^;{]*\)\s*\{", text, re.M)
    assert m, name
    i, depth = m.end(), 1
    start = i
    while depth:
        depth += {"{": 1, "}": -1}.get(text[i], 0)
        i += 1
    body = re.sub(r"//[^\n]*", "", text[start:i])
    return sorted(float(x.group(0).rstrip("f")) for x in LITERAL.finditer(body))


@pytest.mark.parametrize("source", ["oracle/oracle_bsdf.hpp", "tinytrace_amd/csrc/device_bsdf.h"])
def test_fit_constants_equal_the_reference(source):
    fixture = json.load(open(os.path.join(ROOT, "tests", "golden", "bsdf_fit_constants.json")))
    text = open(os.path.join(ROOT, source)).read()
    for name, entry in fixture.items():
        if name == "energy_compensated_oren_nayar" and "device" in source:
            continue  # restructured there (hoisted wi term); GPU parity pins it to the oracle bit for bit
        want = sorted(entry["literals"])
        if name == "oren_nayar":
            want.remove(3.0)  # the reference's pow(s, 3.0) is written s * s * s
        assert function_literals(text, name) == want, (source, name, entry["source"])


# ---------------------------------------------------------- planted slips --
# (pin, exact text in oracle_bsdf.hpp, its replacement)
SLIPS = [
    (check_fresnel_metal, "c * pow6(R(1) - c)", "c * pow5(R(1) - c)"),
    (check_specular_ior, "m.specular_ior / mixs(R(1), m.coat_ior, m.coat_weight)",
     "m.specular_ior / mixs(m.coat_ior, R(1), m.coat_weight)"),
    (check_specular_ior_ratio, "R eps = rsign(ior - R(1))", "R eps = rsign(R(1) - ior)"),
    (check_specular_fresnel_with_coat, "(R(1) - pow2(c)) / pow2(m.coat_ior)", "(R(1) - pow2(c)) / m.coat_ior"),
    (check_average_dielectric_fresnel, "R(1) - pow2(ior) * (R(1) - dieletric_fresnel_factor",
     "R(1) - ior * (R(1) - dieletric_fresnel_factor"),
    (check_oren_nayar, "(R(1) - s * s * s)", "(R(1) - s * s)"),
    (check_energy_compensated_oren_nayar, "V3<R> rho_ms = rho * rho * average_albedo",
     "V3<R> rho_ms = rho * average_albedo"),
    (check_fuzz_ltc, "R jacobian = pow2(a_inverse / length_squared(direction));",
     "R jacobian = a_inverse / length_squared(direction);"),
    (check_fuzz_ltc, "R albedo = fuzz_albedo(cos_theta(wi), roughness);", "R albedo = fuzz_albedo(roughness, cos_theta(wi));"),
    (check_smooth_limits, "R(1) - fresnel_dielectric(rabs(dot(wir, mn)), R(1) / ior)",
     "R(1) - fresnel_dielectric(rabs(dot(wir, mn)), ior)"),
    (check_smooth_limits, "R fresnel = fresnel_dielectric(rabs(d), ior);", "R fresnel = fresnel_dielectric(R(1) - rabs(d), ior);"),
    (check_lobe_weights, "(R(1) - drc) / (R(1) - base_albedo * drc)", "(R(1) - drc) / (R(1) - base_albedo)"),
    (check_lobe_weights, "mixv(V3<R>(R(1)), R(1) - l.albedos[FUZZ], m.fuzz_weight)",
     "mixv(V3<R>(R(1)), l.albedos[FUZZ], m.fuzz_weight)"),
    (check_lobe_weights, "(R(1) - average_dielectric_fresnel(m.coat_ior)) / pow2(m.coat_ior)",
     "(R(1) - average_dielectric_fresnel(m.coat_ior)) / m.coat_ior"),
    (check_lobe_weights, "l.weights[DIFFUSE] = opaque * (R(1) - l.albedos[SPECULAR_BRDF]);",
     "l.weights[DIFFUSE] = opaque;"),
    (check_btdf_refraction, "R jacobian = pow2(ior) * rabs(cos_theta(wi)) /", "R jacobian = ior * rabs(cos_theta(wi)) /"),
    (check_btdf_refraction, "rmax(pow2(cos_theta(wo) + ior * cos_theta(wi)), K<R>::DENOM_TOLERANCE);",
     "rmax(pow2(cos_theta(wi) + ior * cos_theta(wo)), K<R>::DENOM_TOLERANCE);"),
    (check_btdf_refraction, "V3<R> mn = -wor - ior * wir;", "V3<R> mn = -wor - wir / ior;"),
    (check_fuzz_sample, "direction.x / ic.x - direction.z * ic.y / ic.x", "direction.x / ic.x + direction.z * ic.y / ic.x"),
    (check_fuzz_sample, "    ok = true;\n    return fuzz_density_and_brdf(m, direction,",
     "    ok = true;\n    return fuzz_density_and_brdf(m, wn,"),
]


def build_slip(tmp, index, old, new):
    d = os.path.join(tmp, f"slip{index}")
    os.makedirs(d)
    text = open(os.path.join(ORACLE, "oracle_bsdf.hpp")).read()
    assert text.count(old) == 1, old
    open(os.path.join(d, "oracle_bsdf.hpp"), "w").write(text.replace(old, new))
    shutil.copy(os.path.join(ORACLE, "bsdf_hooks.cpp"), d)
    out = os.path.join(d, "hooks.so")
    # The copy's own oracle_bsdf.hpp is found first; the rest from oracle/.
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fno-math-errno",
                    "-I", ORACLE, "-shared", "-o", out, os.path.join(d, "bsdf_hooks.cpp")], check=True)
    return out


def test_planted_slips_fail_their_pins():
    # The last build is the unchanged header: the control that passes all.
    cases = SLIPS + [(None, "template <typename R>\ninline V3<R> fresnel_metal",
                      "template <typename R>\ninline V3<R> fresnel_metal")]
    with tempfile.TemporaryDirectory() as tmp:
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
            libs = list(pool.map(lambda a: build_slip(tmp, a[0], a[1][1], a[1][2]), enumerate(cases)))
        for pin in PINS:
            pin(binding.BsdfHooks(libs[-1], precision=1))
        for (pin, old, new), path in zip(SLIPS, libs):
            with pytest.raises(AssertionError):
                pin(binding.BsdfHooks(path, precision=1))
