"""Oracle BxDFs and lights against closed forms (Tracer/MaterialsDefault.hpp,
DistributionFunctions.h, LightsDefault.hpp).

Pins: Lambert reflectance/pdf identity (white furnace per sample), pdf
normalisation over the sphere, reciprocity of Unreal's Evaluate, sample/pdf
consistency, Snell + Fresnel closed forms for Refract, the mirror law for
Reflect, and -- as Tests/Tracer/T_DefaultLights.cu:322-392 does -- that a light
sample's pdf equals the pdf of hitting that light at the sampled point.
"""
import math

import numpy as np
import pytest

from mray_amd import abi, scenes
from oracle import refcpu

ID = (1.0, 0.0, 0.0, 0.0)          # world == tangent (normal +Z)
NZ = (0.0, 0.0, 1.0)


def _dirs(n, seed, upper=True):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    if upper:
        d[:, 2] = np.abs(d[:, 2])
    return d.astype(np.float32)


LAMBERT = {"type": abi.MAT_LAMBERT, "albedo": (0.8, 0.5, 0.2)}
UNREAL = [{"type": abi.MAT_UNREAL, "albedo": (0.7, 0.4, 0.3), "roughness": r, "metallic": m,
           "specular": 0.5} for r, m in ((0.3, 0.0), (0.6, 1.0), (0.15, 0.4))]


def test_lambert_furnace_per_sample():
    rng = np.random.default_rng(1)
    wo = _dirs(1, 2)[0]
    for _ in range(500):
        rn = rng.uniform(size=4).astype(np.float32)
        out = refcpu.bsdf_probe(LAMBERT, ID, NZ, 0, wo, rn=rn, op=0)
        wi, pdf, refl = out[:3], out[3], out[4:7]
        assert abs(np.linalg.norm(wi) - 1) < 1e-6 and wi[2] >= 0
        assert pdf == pytest.approx(wi[2] / math.pi, rel=1e-5)
        if pdf > 1e-4:
            assert np.allclose(refl / pdf, LAMBERT["albedo"], rtol=1e-4)
            ev = refcpu.bsdf_probe(LAMBERT, ID, NZ, 0, wo, wi=wi, op=1)
            assert np.allclose(ev[:3], refl, rtol=1e-5)
            p = refcpu.bsdf_probe(LAMBERT, ID, NZ, 0, wo, wi=wi, op=2)[0]
            assert p == pytest.approx(pdf, rel=1e-5)


@pytest.mark.parametrize("mat", [LAMBERT] + UNREAL)
def test_pdf_integrates_to_one(mat):
    """Monte Carlo over the full sphere with uniform directions."""
    wo = np.array([0.3, -0.2, 0.93], np.float32)
    wo /= np.linalg.norm(wo)
    d = _dirs(60000, 3, upper=False)
    pdfs = np.array([refcpu.bsdf_probe(mat, ID, NZ, 0, wo, wi=w, op=2)[0] for w in d])
    integral = pdfs.mean() * 4 * math.pi
    assert integral == pytest.approx(1.0, abs=0.03)


@pytest.mark.parametrize("mat", UNREAL)
def test_unreal_reciprocity(mat):
    """f(wi, wo) = f(wo, wi); the reflectance carries cos(wi), so
    refl(wi, wo) * cos(wo) = refl(wo, wi) * cos(wi)."""
    a = _dirs(300, 4)
    b = _dirs(300, 5)
    for wi, wo in zip(a, b):
        r1 = refcpu.bsdf_probe(mat, ID, NZ, 0, wo, wi=wi, op=1)[:3] * wo[2]
        r2 = refcpu.bsdf_probe(mat, ID, NZ, 0, wi, wi=wo, op=1)[:3] * wi[2]
        assert np.allclose(r1, r2, rtol=2e-4, atol=1e-7)


@pytest.mark.parametrize("mat", UNREAL)
def test_unreal_sample_pdf_consistency(mat):
    rng = np.random.default_rng(7)
    wo = np.array([0.1, 0.4, 0.91], np.float32)
    wo /= np.linalg.norm(wo)
    checked = 0
    for _ in range(400):
        rn = rng.uniform(size=4).astype(np.float32)
        out = refcpu.bsdf_probe(mat, ID, NZ, 0, wo, rn=rn, op=0)
        wi, pdf, refl = out[:3], out[3], out[4:7]
        if wi[2] <= 1e-3 or pdf <= 1e-3:
            continue
        p = refcpu.bsdf_probe(mat, ID, NZ, 0, wo, wi=wi, op=2)[0]
        ev = refcpu.bsdf_probe(mat, ID, NZ, 0, wo, wi=wi, op=1)[:3]
        assert p == pytest.approx(pdf, rel=2e-3)
        assert np.allclose(ev, refl, rtol=2e-3, atol=1e-6)
        checked += 1
    assert checked > 200


def test_reflect_mirror_law():
    m = {"type": abi.MAT_REFLECT}
    for wo in _dirs(50, 8):
        out = refcpu.bsdf_probe(m, ID, NZ, 0, wo, rn=(0, 0, 0, 0), op=0)
        assert np.allclose(out[:3], [-wo[0], -wo[1], wo[2]], atol=1e-6)
        assert out[3] == 1.0 and np.allclose(out[4:8], 1.0)


def test_refract_snell_and_fresnel():
    m = {"type": abi.MAT_REFRACT, "cauchyFront": (1.0, 0, 0), "cauchyBack": (1.5, 0, 0)}
    # normal incidence: F = ((1 - 1.5) / (1 + 1.5))^2 = 0.04
    wo = np.array([0, 0, 1], np.float32)
    r_lo = refcpu.bsdf_probe(m, ID, NZ, 0, wo, rn=(0.039, 0, 0, 0), op=0)
    r_hi = refcpu.bsdf_probe(m, ID, NZ, 0, wo, rn=(0.041, 0, 0, 0), op=0)
    assert r_lo[3] == pytest.approx(0.04, rel=1e-5)          # reflected with pdf F
    assert np.allclose(r_lo[:3], [0, 0, 1], atol=1e-6)
    assert r_hi[3] == pytest.approx(0.96, rel=1e-5)          # transmitted, pdf 1 - F
    assert np.allclose(r_hi[:3], [0, 0, -1], atol=1e-6)
    # oblique: sin(t) = sin(i) / 1.5, throughput refl / pdf == 1
    for wo in _dirs(40, 9):
        out = refcpu.bsdf_probe(m, ID, NZ, 0, wo, rn=(0.999, 0, 0, 0), op=0)
        wi = out[:3]
        sin_i = math.sqrt(max(0.0, 1 - wo[2] ** 2))
        sin_t = math.sqrt(max(0.0, 1 - wi[2] ** 2))
        assert wi[2] < 0
        assert sin_t == pytest.approx(sin_i / 1.5, abs=2e-6)
        assert np.allclose(out[4:7] / out[3], 1.0, rtol=1e-6)


def test_triangle_light_sample_pdf_matches_hit_pdf():
    """LightPrim::SampleSolidAngle pdf == LightPrim::PdfSolidAngle at the sampled
    point (T_DefaultLights.cu:322-392 pattern), through the uniform sampler."""
    s = scenes.cornell()
    o = refcpu.OracleScene(s)
    rng = np.random.default_rng(10)
    light_count = len(s.light_tris) + 1
    checked = 0
    for _ in range(2000):
        p = rng.uniform(-0.9, 0.9, size=3).astype(np.float32)
        rn = rng.uniform(size=3).astype(np.float32)
        out = o.light_probe(p, rn)
        li = min(int(rn[2] * light_count), light_count - 1)
        if li >= len(s.light_tris) or out[3] == 0.0:
            continue
        q = out[:3]
        d = (q - p) / np.linalg.norm(q - p)
        tri = s.positions[s.indices[s.light_tris[li]]].astype(np.float64)
        # barycentrics (a, b) of q for MetaHit: q = p0 a + p1 b + p2 (1 - a - b)
        A = np.stack([tri[0] - tri[2], tri[1] - tri[2]], 1)
        ab, *_ = np.linalg.lstsq(A, q.astype(np.float64) - tri[2], rcond=None)
        pdf_hit = o.light_pdf(li, ab[0], ab[1], p, d)
        assert pdf_hit == pytest.approx(float(out[3]), rel=1e-4)
        checked += 1
    assert checked > 500


def test_null_boundary_light_contributes_nothing():
    s = scenes.cornell()
    o = refcpu.OracleScene(s)
    out = o.light_probe(np.zeros(3, np.float32), np.array([0.3, 0.3, 0.99], np.float32))
    assert out[3] == 0.0 and np.all(out[4:8] == 0.0)
