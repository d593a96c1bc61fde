"""Independent pins for oracle functions that the image tests cannot see.

Each check compares the oracle with mathematics it does not share code with:

* InvErf (the Gaussian film filter's sampler, Filters.h:195-220 over
  Math::InvErrFunc, Core/Math.h:787-833; the single-precision polynomial of
  M. Giles, "Approximating the erfinv function", GPU Computing Gems, 2010)
  against scipy.special.erfinv in double;
* SampleGaussian (DistributionFunctions.h:686-705) against the Gaussian CDF
  and density, including the +-3.5 sigma clamp at xi = 0;
* the pinhole camera ray (CamerasDefault.hpp:8-36, 97-151 with
  RayGenKernels.kt.h:182-225) against a closed-form construction in double:
  camera basis from position / gaze / up, image plane at the near distance,
  pixel (i mod W, i div W) from the bottom-left, jitter 0.5 + the filter
  offset;
* GGX (Walter et al. 2007), Smith's height-correlated masking-shadowing
  (Heitz 2014), Schlick's Fresnel and Burley's diffuse (2012) as UE4's
  model combines them (Karis 2013): the oracle's Unreal BxDF pointwise
  against a textbook numpy evaluation, and the weak white furnace
  (integral of D G1 / (4 cos_o) over the sphere = 1, Heitz 2014) for the
  textbook terms, so a constant-factor mistake in D F G / (4 NdV) fails;
* the latitude-longitude Skysphere map (LightsDefault.hpp:177-222,
  GraphicsFunctions.h:253-285) against closed-form directions, its round
  trip, and its solid-angle pdf integrating to one over the sphere.
"""
import math

import numpy as np
import pytest
from scipy import special, stats

from mray_amd import abi, scenes
from oracle import refcpu

ID = (1.0, 0.0, 0.0, 0.0)          # world == tangent (normal +Z)
NZ = (0.0, 0.0, 1.0)


# ---------------------------------------------------------------------------
# Inverse error function and the Gaussian filter sampler
# ---------------------------------------------------------------------------
def test_inv_erf_against_scipy():
    x = np.linspace(-0.9999995, 0.9999995, 1_000_001).astype(np.float32)
    got = refcpu.inv_erf(x).astype(np.float64)
    want = special.erfinv(x.astype(np.float64))
    err = np.abs(got - want)
    ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    # measured: 2.3 ulp at most; a wrong coefficient moves this by orders
    assert (err / ulp).max() < 4.0
    nz = np.abs(want) > 1e-6
    assert (err[nz] / np.abs(want[nz])).max() < 1e-6
    # odd symmetry and the exact zero
    assert np.array_equal(refcpu.inv_erf(-x[::-1]), -got[::-1].astype(np.float32))
    assert refcpu.inv_erf(np.zeros(1))[0] == 0.0


def test_sample_gaussian_is_the_inverse_cdf():
    sigma = 0.285714
    n = 200_001
    xi = ((np.arange(n) + 0.5) / n).astype(np.float32)
    v, pdf = refcpu.sample_gaussian(xi, sigma)
    v64 = v.astype(np.float64)
    assert np.all(np.diff(v64) >= 0.0)
    # Phi(v / sigma) = xi up to float rounding of xi and v
    cdf = stats.norm.cdf(v64 / sigma)
    assert np.abs(cdf - xi.astype(np.float64)).max() < 3e-6
    # pdf is the Gaussian density at the sample
    assert np.allclose(pdf, stats.norm.pdf(v64, scale=sigma), rtol=2e-5)
    # stratified samples reproduce the distribution: KS distance ~ 1/n
    ks = stats.kstest(v64 / sigma, "norm").statistic
    assert ks < 2e-5
    # xi = 0 gives erfinv(-1) = -inf, clamped to -3.5 sigma
    v0, p0 = refcpu.sample_gaussian(np.zeros(1), sigma)
    assert v0[0] == pytest.approx(-3.5 * sigma, rel=1e-6)
    assert p0[0] == pytest.approx(stats.norm.pdf(3.5, scale=1.0) / sigma, rel=1e-4)


def test_gaussian_sampler_planted_mistake_fails():
    """The CDF check above is sensitive: sigma off by 1 % breaks it."""
    sigma = 0.285714
    xi = ((np.arange(20001) + 0.5) / 20001).astype(np.float32)
    v, _ = refcpu.sample_gaussian(xi, sigma * 1.01)
    cdf = stats.norm.cdf(v.astype(np.float64) / sigma)
    assert np.abs(cdf - xi).max() > 1e-3


# ---------------------------------------------------------------------------
# Pinhole camera rays
# ---------------------------------------------------------------------------
def _closed_form_rays(cam, W, H, xi, filter_radius):
    pos = np.array(cam["position"], np.float64)
    g = np.array(cam["gaze"], np.float64) - pos
    g /= np.linalg.norm(g)
    r = np.cross(g, np.array(cam["up"], np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, g)
    near = cam["nearFar"][0]
    hw = math.tan(cam["fov"][0] / 2) * near
    hh = math.tan(cam["fov"][1] / 2) * near
    sigma = 0.285714 * filter_radius
    off = math.sqrt(2.0) * sigma * special.erfinv(2.0 * xi.astype(np.float64) - 1.0)
    i = np.arange(len(xi))
    px, py = i % W, i // W
    sx = -hw + (px + 0.5 + off[:, 0]) * (2 * hw / W)
    sy = -hh + (py + 0.5 + off[:, 1]) * (2 * hh / H)
    d = g[None] * near + sx[:, None] * r[None] + sy[:, None] * u[None]
    return pos, d / np.linalg.norm(d, axis=1, keepdims=True)


@pytest.mark.parametrize("W,H,filter_radius", [(7, 5, 1.0), (16, 16, 2.0)])
def test_pinhole_rays_closed_form(W, H, filter_radius):
    sc = scenes.cornell(aspect=W / H)
    sc.camera["position"] = (0.3, 0.2, 3.5)
    sc.camera["gaze"] = (-0.1, 0.05, 0.0)
    params = abi.render_params(W, H, 1, seed=7, filter_radius=filter_radius)
    o = refcpu.OracleScene(sc)
    rays, weights = o.camera_rays(params, W * H)
    o.close()
    # the slot's first two PCG32 draws (the RNG itself is pinned in test_oracle_rng)
    states = refcpu.seed_states(7, W * H)
    xi = np.array([[refcpu.to_float01(v) for v in refcpu.pcg_sequence(int(s), 2)] for s in states],
                  np.float64)
    pos, d = _closed_form_rays(sc.camera, W, H, xi, filter_radius)
    assert np.allclose(rays[:, 0:3], pos[None], atol=0)
    assert np.allclose(rays[:, 3], sc.camera["nearFar"][0])
    assert np.allclose(rays[:, 7], sc.camera["nearFar"][1])
    assert np.abs(rays[:, 4:7] - d).max() < 3e-6
    # weight = filter value / sample pdf = 1 for the Gaussian's own samples
    assert np.allclose(weights, 1.0, rtol=1e-5)


def test_pinhole_orientation():
    """Slot 0 is the bottom-left pixel; the top edge ray differs from the
    bottom edge ray by fovY."""
    W = H = 64
    sc = scenes.cornell()
    o = refcpu.OracleScene(sc)
    rays, _ = o.camera_rays(abi.render_params(W, H, 1, seed=3, filter_radius=1e-4), W * H)
    o.close()
    d = rays[:, 4:7].reshape(H, W, 3)
    assert d[0, 0, 0] < 0 and d[0, 0, 1] < 0          # bottom-left: -x, -y
    assert d[-1, -1, 0] > 0 and d[-1, -1, 1] > 0      # top-right: +x, +y
    bottom = d[0, W // 2 - 1] + d[0, W // 2]
    top = d[-1, W // 2 - 1] + d[-1, W // 2]
    ang = math.acos(np.dot(bottom, top) / np.linalg.norm(bottom) / np.linalg.norm(top))
    fov_y = sc.camera["fov"][1]
    # pixel centres span (H - 1) / H of the plane
    want = 2 * math.atan(math.tan(fov_y / 2) * (H - 1) / H)
    assert ang == pytest.approx(want, rel=2e-4)


# ---------------------------------------------------------------------------
# GGX / Smith / Schlick / Burley
# ---------------------------------------------------------------------------
def _lambda(c, a):
    c2 = np.minimum(c * c, 1.0)
    tan2 = (1.0 - c2) / c2
    return 0.5 * (np.sqrt(1.0 + a * a * tan2) - 1.0)


def _ggx_d(nh, a):
    return a * a / (math.pi * (nh * nh * (a * a - 1.0) + 1.0) ** 2)


def _unreal_textbook(mat, wo, wi):
    """f * cos_i of the UE4 model in double: Burley diffuse + GGX specular."""
    rough, metal, spec = mat["roughness"], mat["metallic"], mat["specular"]
    alb = np.array(mat["albedo"], np.float64)
    a = rough * rough
    h = wo[None] + wi
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    nl, nv, nh = wi[:, 2], wo[2], h[:, 2]
    vh = h @ wo
    lh = np.sum(wi * h, axis=1)
    f0 = 0.08 * spec * (1 - metal) + alb * metal
    F = f0[None] + (1 - f0[None]) * (1 - vh)[:, None] ** 5
    G2 = 1.0 / (1.0 + _lambda(nv, a) + _lambda(nl, a))
    specular = (_ggx_d(nh, a) * G2 / (4 * nv * nl))[:, None] * F
    fd90 = 0.5 + 2 * rough * lh * lh
    burley = (1 + (fd90 - 1) * (1 - nl) ** 5) * (1 + (fd90 - 1) * (1 - nv) ** 5)
    diffuse = ((1 - metal) / math.pi * burley)[:, None] * alb[None]
    return (diffuse + specular) * nl[:, None]


UNREAL = [{"type": abi.MAT_UNREAL, "albedo": (0.7, 0.4, 0.3), "roughness": r, "metallic": m,
           "specular": s} for r, m, s in ((0.3, 0.0, 0.5), (0.6, 1.0, 0.5), (0.15, 0.4, 0.8),
                                          (0.9, 0.2, 0.0))]


@pytest.mark.parametrize("mat", UNREAL)
def test_unreal_matches_textbook(mat):
    rng = np.random.default_rng(11)
    for cos_o in (1.0, 0.8, 0.35):
        wo = np.array([math.sqrt(1 - cos_o ** 2), 0.0, cos_o])
        wi = rng.normal(size=(4000, 3))
        wi /= np.linalg.norm(wi, axis=1, keepdims=True)
        wi[:, 2] = np.abs(wi[:, 2]) + 0.02
        wi /= np.linalg.norm(wi, axis=1, keepdims=True)
        got, _ = refcpu.bsdf_eval_many(mat, ID, NZ, 0, wo, wi)
        want = _unreal_textbook(mat, wo, wi)
        scale = np.maximum(np.abs(want), 1e-3)
        assert (np.abs(got[:, :3] - want) / scale).max() < 2e-3


def _half_vector_grid(a, n_t=3000, n_p=256):
    """Half vectors on a grid that follows the GGX lobe: theta_h =
    atan(a tan(s pi / 2)); weights are solid-angle quadrature weights."""
    s = (np.arange(n_t) + 0.5) / n_t
    th = np.arctan(a * np.tan(s * math.pi / 2))
    dth = a * (math.pi / 2) / np.cos(s * math.pi / 2) ** 2 / (1 + (a * np.tan(s * math.pi / 2)) ** 2) / n_t
    ph = (np.arange(n_p) + 0.5) / n_p * 2 * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    h = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    w = (np.sin(T) * dth[:, None] * (2 * math.pi / n_p)).ravel()
    return h, w


@pytest.mark.parametrize("rough", [0.3, 0.6])
@pytest.mark.parametrize("cos_o", [1.0, 0.5])
def test_weak_white_furnace(rough, cos_o):
    """Heitz 2014: the integral over wi of D(h) G1(wo) / (4 cos_o) is one
    (F = 1).  In half-vector space dwi = 4 |wo.h| dh, so the integrand is
    D(h) G1(wo) |wo.h| / cos_o over h with wo.h > 0.  The oracle's metallic
    white Unreal (F0 = 1, diffuse 0) returns D G2 / (4 cos_o) per unit
    solid angle, which the same quadrature must bring to the textbook
    integral with G2 (below one by the masking-shadowing loss)."""
    a = rough * rough
    wo = np.array([math.sqrt(1 - cos_o ** 2), 0.0, cos_o])
    h, w = _half_vector_grid(a)
    oh = h @ wo
    keep = oh > 0
    h, w, oh = h[keep], w[keep], oh[keep]
    g1 = 1.0 / (1.0 + _lambda(cos_o, a))
    weak = np.sum(_ggx_d(h[:, 2], a) * g1 * oh / cos_o * w)
    assert weak == pytest.approx(1.0, abs=2e-3)
    wi = 2 * oh[:, None] * h - wo[None]
    up = wi[:, 2] > 1e-6
    mat = {"type": abi.MAT_UNREAL, "albedo": (1.0, 1.0, 1.0), "roughness": rough,
           "metallic": 1.0, "specular": 0.5}
    got, _ = refcpu.bsdf_eval_many(mat, ID, NZ, 0, wo, wi[up])
    jac = 4 * oh[up]
    e_oracle = np.sum(got[:, 0].astype(np.float64) * jac * w[up])
    g2 = 1.0 / (1.0 + _lambda(cos_o, a) + _lambda(wi[up, 2], a))
    e_text = np.sum(_ggx_d(h[up, 2], a) * g2 / (4 * cos_o) * jac * w[up])
    assert e_oracle == pytest.approx(e_text, rel=2e-3)
    assert 0.7 < e_oracle <= 1.0 + 2e-3


# ---------------------------------------------------------------------------
# Latitude-longitude Skysphere map
# ---------------------------------------------------------------------------
def test_spherical_map_closed_form():
    # u = (atan2(y, x) + pi) / (2 pi), v = 1 - acos(z) / pi  (Z-up)
    cases = {(0.5, 0.5): (1.0, 0.0, 0.0), (0.75, 0.5): (0.0, 1.0, 0.0),
             (0.25, 0.5): (0.0, -1.0, 0.0), (0.0, 0.5): (-1.0, 0.0, 0.0)}
    uv = np.array(list(cases.keys()), np.float32)
    d, s = refcpu.spherical_to_dir(uv)
    assert np.allclose(d, np.array(list(cases.values())), atol=1e-6)
    assert np.allclose(s, 1.0, atol=1e-6)
    # v = 1 is the +Z pole (Y-up: +Y), v = 0 the -Z pole
    d, s = refcpu.spherical_to_dir(np.array([[0.3, 1.0], [0.8, 0.0]], np.float32))
    assert np.allclose(d, [[0, 0, 1], [0, 0, -1]], atol=1e-6)
    assert np.allclose(s, 0.0, atol=1e-6)
    back = refcpu.dir_to_spherical(np.array(list(cases.values()), np.float32))
    # -X sits on the seam: atan2(+0, -1) = pi gives u = 1, the same texel column as u = 0
    du = np.abs(back[:, 0] - uv[:, 0])
    assert np.all(np.minimum(du, 1 - du) < 1e-6)
    assert np.allclose(back[:, 1], uv[:, 1], atol=1e-6)


def test_spherical_round_trip_and_pdf():
    rng = np.random.default_rng(5)
    uv = rng.uniform(0.01, 0.99, size=(20000, 2)).astype(np.float32)
    d, s = refcpu.spherical_to_dir(uv)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert np.allclose(s, np.sqrt(np.maximum(0, 1 - d[:, 2].astype(np.float64) ** 2)), atol=1e-5)
    assert np.abs(refcpu.dir_to_spherical(d) - uv).max() < 2e-5
    # closed form, independently: theta = 2 pi u - pi, inclination = (1 - v) pi
    th = 2 * math.pi * uv[:, 0].astype(np.float64) - math.pi
    inc = (1 - uv[:, 1].astype(np.float64)) * math.pi
    want = np.stack([np.cos(th) * np.sin(inc), np.sin(th) * np.sin(inc), np.cos(inc)], -1)
    assert np.abs(d - want).max() < 2e-6
    # a uniform uv density (pdf 1) becomes a solid-angle pdf that integrates
    # to one over the sphere: midpoint rule in (u, v), dw = 2 pi^2 sin du dv
    n = 400
    g = (np.arange(n) + 0.5) / n
    U, V = np.meshgrid(g, g, indexing="ij")
    _, sg = refcpu.spherical_to_dir(np.stack([U.ravel(), V.ravel()], -1))
    pdf = np.array([refcpu.sky_solid_angle_pdf(abi.BOUNDARY_SKYSPHERE_SPHERICAL, 1.0, float(x))
                    for x in sg[: n * 4]])
    # pdf * dw = du dv, exactly one cell's area, wherever sin > 0
    dw = 2 * math.pi ** 2 * sg[: n * 4].astype(np.float64) / (n * n)
    assert np.allclose(pdf * dw, 1.0 / (n * n), rtol=1e-5)
    sphere = np.sum(2 * math.pi ** 2 * sg.astype(np.float64) / (n * n))
    assert sphere == pytest.approx(4 * math.pi, rel=1e-4)
    # the pole has zero pdf, as the reference's guard gives
    assert refcpu.sky_solid_angle_pdf(abi.BOUNDARY_SKYSPHERE_SPHERICAL, 1.0, 0.0) == 0.0
