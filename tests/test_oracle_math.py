"""Pins the oracle's scalar math and shading functions against the
mathematics (closed forms in double precision, libm, IEC 61966-2-1) and the
reference's one known-answer test."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_host as oh
from rtg import abi

L = oh.lib()
fp = C.POINTER(C.c_float)


def f32(x):
    return float(np.float32(x))


def ulp_diff(a, b):
    a = np.float32(a)
    b = np.float32(b)
    ia = np.array([a]).view(np.int32)[0]
    ib = np.array([b]).view(np.int32)[0]
    if ia < 0:
        ia = np.int32(-2147483648) - ia
    if ib < 0:
        ib = np.int32(-2147483648) - ib
    return abs(int(ia) - int(ib))


# ---------------------------------------------------------------------------
# The reference's known-answer test: shaders/ray-tracing/src/pbr.rs:52-69
# ---------------------------------------------------------------------------
def test_ggx_division_by_zero():
    # view (1,0,0), normal (0,1,0): NoV = 0 clamps to the epsilon; roughness 0
    # (pbr.glsl clamps NoV to 10.0e-10 = 1e-9)
    nov = f32(1e-9)
    v = L.orc_v_smith_ggx_correlated(nov, 1.0, 0.0)
    assert math.isfinite(v)
    # closed form at roughness 0: 0.5 / (NoL*NoV + NoV*NoL) = 0.25 / NoV
    assert v == pytest.approx(0.25 / nov, rel=1e-6)


def test_v_smith_ggx_matches_filament_in_double():
    rng = np.random.default_rng(1)
    for _ in range(200):
        nov, nol = rng.uniform(1e-3, 1.0, 2)
        a = rng.uniform(0.01, 1.0)
        nov, nol, a = f32(nov), f32(nol), f32(a)
        a2 = a * a
        exact = 0.5 / (nol * math.sqrt(nov * nov * (1 - a2) + a2) + nov * math.sqrt(nol * nol * (1 - a2) + a2))
        assert L.orc_v_smith_ggx_correlated(nov, nol, a) == pytest.approx(exact, rel=2e-6)


def test_d_ggx_matches_closed_form():
    for noh in (0.0, 0.3, 0.7, 0.999, 1.0):
        for a in (0.05, 0.3, 1.0):
            noh, a = f32(noh), f32(a)
            den = 1 - noh * noh + (noh * a) ** 2
            exact = (a / den) ** 2 / math.pi
            # 1 - NoH^2 cancels near NoH = 1: the FP32 error grows with NoH^2 / den
            cond = 1 + noh * noh / den
            assert L.orc_d_ggx(noh, a) == pytest.approx(exact, rel=4e-7 * 4 * cond)
    # alpha = 1 makes D uniform: 1/pi for every NoH
    assert L.orc_d_ggx(0.3, 1.0) == pytest.approx(1 / math.pi, rel=1e-6)


def brdf_double(n, v, l, pr, base, metallic, reflectance, intensity):
    n, v, l, base = (np.asarray(x, dtype=np.float64) for x in (n, v, l, base))
    h = (v + l) / np.linalg.norm(v + l)
    a = pr * pr
    nov = min(max(n @ v, 1e-9), 1.0)
    noh = min(max(n @ h, 0.0), 1.0)
    nol = min(max(n @ l, 0.0), 1.0)
    loh = min(max(l @ h, 0.0), 1.0)
    D = (a / (1 - noh * noh + (noh * a) ** 2)) ** 2 / math.pi
    f0 = 0.16 * reflectance ** 2 * (1 - metallic) + base * metallic
    f90 = 0.5 + 2 * a * loh * loh  # Burley f90, also the specular f90 (pbr.glsl:194-196)
    F = f0 + (f90 - f0) * (1 - loh) ** 5
    a2 = a * a
    V = 0.5 / (nol * math.sqrt(nov * nov * (1 - a2) + a2) + nov * math.sqrt(nol * nol * (1 - a2) + a2))
    fd = (1 + (f90 - 1) * (1 - nol) ** 5) * (1 + (f90 - 1) * (1 - nov) ** 5) / math.pi
    return intensity * nol * (base * fd + D * V * F)


def test_brdf_matches_filament_in_double():
    rng = np.random.default_rng(2)
    for _ in range(300):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if n @ v < 0:
            v = -v
        l = rng.normal(size=3)
        l /= np.linalg.norm(l)
        pr = rng.uniform(0.2, 1.0)
        base = rng.uniform(0, 1, 3)
        m = rng.uniform()
        nf, vf, lf, bf = (np.ascontiguousarray(x, dtype=np.float32) for x in (n, v, l, base))
        out = np.zeros(3, dtype=np.float32)
        L.orc_brdf(oh.fptr(nf), oh.fptr(vf), oh.fptr(lf), f32(pr), oh.fptr(bf), f32(m), 0.5, 0.75, oh.fptr(out))
        ref = brdf_double(nf.astype(np.float64), vf.astype(np.float64), lf.astype(np.float64), f32(pr),
                          bf.astype(np.float64), f32(m), 0.5, 0.75)
        np.testing.assert_allclose(out, ref, rtol=2e-4, atol=2e-6)


def test_brdf_dark_below_horizon():
    n = np.array([0, 1, 0], np.float32)
    v = np.array([0, 1, 0], np.float32)
    l = np.array([0, -1, 0], np.float32)
    base = np.array([0.5, 0.5, 0.5], np.float32)
    out = np.ones(3, np.float32)
    L.orc_brdf(oh.fptr(n), oh.fptr(v), oh.fptr(l), 0.5, oh.fptr(base), 0.0, 0.5, 1.0, oh.fptr(out))
    assert np.all(out == 0)


# ---------------------------------------------------------------------------
# Own transcendentals vs libm in double
# ---------------------------------------------------------------------------
def test_sin_cos_against_libm():
    xs = np.linspace(0.0, 2 * math.pi, 20001, dtype=np.float32)
    worst = 0
    for x in xs:
        c = L.orc_cos(float(x))
        s = L.orc_sin(float(x))
        rc, rs = math.cos(float(x)), math.sin(float(x))
        assert abs(c - rc) <= 1.2e-7 * max(1.0, abs(rc)) + 1e-8
        assert abs(s - rs) <= 1.2e-7 * max(1.0, abs(rs)) + 1e-8
        worst = max(worst, ulp_diff(c, rc) if abs(rc) > 1e-3 else 0)
    assert worst <= 2
    assert L.orc_cos(0.0) == 1.0 and L.orc_sin(0.0) == 0.0


def test_log2_exp2_pow_against_libm():
    # log2 feeds pow through y * log2(x): bound its absolute error (2^-22 scale)
    for x in np.geomspace(1e-30, 1e30, 4001, dtype=np.float32):
        r = math.log2(float(x))
        assert abs(L.orc_log2(float(x)) - r) <= 2.4e-7 * max(1.0, abs(r))
    for y in np.linspace(-140.0, 127.0, 4001, dtype=np.float32):
        assert ulp_diff(L.orc_exp2(float(y)), 2.0 ** float(y)) <= 2
    for x in np.linspace(0.0031308, 4.0, 2001, dtype=np.float32):
        r = float(x) ** (1 / 2.4)
        assert abs(L.orc_pow(float(x), f32(1 / 2.4)) - r) <= 4e-7 * r
    assert L.orc_log2(1.0) == 0.0
    assert L.orc_exp2(0.0) == 1.0
    assert L.orc_exp2(-150.5) == 0.0
    assert math.isinf(L.orc_exp2(128.0))
    assert L.orc_pow(0.0, 0.4) == 0.0
    assert L.orc_exp2(-149.0) == np.float32(2.0 ** -149)  # smallest subnormal, exact


def test_linear_to_srgb_matches_iec():
    xs = np.concatenate([np.linspace(0, 0.0031308, 50), np.linspace(0.0031309, 1.0, 500)]).astype(np.float32)
    inp = np.zeros(3, np.float32)
    out = np.zeros(3, np.float32)
    for x in xs:
        inp[:] = x
        L.orc_linear_to_srgb(oh.fptr(inp), oh.fptr(out))
        ref = 12.92 * x if x <= 0.0031308 else 1.055 * float(x) ** (1 / 2.4) - 0.055
        assert abs(out[0] - ref) <= 1e-6


def test_srgb_decode_table():
    lut = np.zeros(256, np.float32)
    L.orc_srgb_lut(oh.fptr(lut))
    for i in (0, 1, 10, 11, 128, 255):
        c = i / 255
        ref = c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4
        assert lut[i] == np.float32(ref)
    assert lut[255] == 1.0 and np.all(np.diff(lut) > 0)


# ---------------------------------------------------------------------------
# Light sampling (closest_hit_textured.glsl:77-94) and the heatmap ramp
# ---------------------------------------------------------------------------
def test_sample_directional_light_cone():
    rng = np.random.default_rng(3)
    centre = np.array([0.3, 0.8, 0.52], np.float64)
    centre /= np.linalg.norm(centre)
    c32 = centre.astype(np.float32)
    radius = f32(math.tan(math.radians(1.0)))
    out = np.zeros(3, np.float32)
    for _ in range(500):
        r = rng.uniform(0, 1, 2).astype(np.float32)
        L.orc_sample_directional_light(oh.fptr(r), oh.fptr(c32), radius, oh.fptr(out))
        assert abs(np.linalg.norm(out) - 1) < 1e-6
        ang = math.acos(min(1.0, float(out.astype(np.float64) @ c32.astype(np.float64))))
        assert ang <= math.atan(radius) * (1 + 1e-4) + 1e-6
    # zero radius: the direction is the (normalised) centre
    L.orc_sample_directional_light(oh.fptr(np.array([0.3, 0.6], np.float32)), oh.fptr(c32), 0.0, oh.fptr(out))
    np.testing.assert_allclose(out, c32, atol=1e-7)
    # rng.x = 1, rng.y = 0 reaches the rim along the tangent
    L.orc_sample_directional_light(oh.fptr(np.array([1.0, 0.0], np.float32)), oh.fptr(c32), radius, oh.fptr(out))
    ang = math.acos(min(1.0, float(out.astype(np.float64) @ c32.astype(np.float64))))
    assert ang == pytest.approx(math.atan(radius), rel=1e-3)


def test_heatmap_ramp_endpoints():
    out = np.zeros(3, np.float32)
    L.orc_heatmap_temperature(0.0, oh.fptr(out))
    # heat 0: wp = 1 - smoothstep(-0.8, 0.8, 0) = 0.5, wc = 0.5 * (1 - 0.5) = 0.25, wn = 0.5
    first, second = np.array([0, 2, 91]) / 255, np.array([0, 108, 251]) / 255
    np.testing.assert_allclose(out, np.clip(0.25 * first + 0.5 * first + 0.5 * second, 0, 1), atol=1e-6)
    # the ramp at cell centres and edges: tests/test_oracle_stages.py


# ---------------------------------------------------------------------------
# Software texture sampling
# ---------------------------------------------------------------------------
def _tex(img, fmt, filt):
    img = np.ascontiguousarray(img)
    t = abi.Texture()
    t.data = img.ctypes.data
    t.height, t.width = img.shape[:2]
    t.format = fmt
    t.filter = filt
    return t, img


def test_nearest_sampling_is_exact_with_repeat():
    img = np.zeros((4, 8, 4), np.uint8)
    img[..., 0] = np.arange(32).reshape(4, 8) * 7
    img[..., 3] = 255
    t, keep = _tex(img, abi.RTG_FORMAT_RGBA8_UNORM, abi.RTG_FILTER_NEAREST)
    lut = np.zeros(256, np.float32)
    out = np.zeros(4, np.float32)
    for (u, v, x, y) in [(0.0, 0.0, 0, 0), (0.99, 0.99, 7, 3), (1.0, 0.0, 0, 0), (-0.01, 0.0, 7, 0),
                         (2.5 / 8, 1.5 / 4, 2, 1), (3.0 / 8, 0.0, 3, 0)]:
        L.orc_sample_texture(C.byref(t), oh.fptr(lut), u, v, oh.fptr(out))
        assert out[0] == np.float32(img[y, x, 0]) / np.float32(255), (u, v)


def test_linear_sampling_weights():
    img = np.zeros((2, 2, 4), np.float32)
    img[0, 0, 0], img[0, 1, 0], img[1, 0, 0], img[1, 1, 0] = 0.0, 1.0, 2.0, 3.0
    t, keep = _tex(img, abi.RTG_FORMAT_RGBA32_FLOAT, abi.RTG_FILTER_LINEAR)
    lut = np.zeros(256, np.float32)
    out = np.zeros(4, np.float32)
    # at texel centres the value is the texel
    L.orc_sample_texture(C.byref(t), oh.fptr(lut), 0.25, 0.25, oh.fptr(out))
    assert out[0] == 0.0
    L.orc_sample_texture(C.byref(t), oh.fptr(lut), 0.75, 0.75, oh.fptr(out))
    assert out[0] == 3.0
    # halfway between the four centres: the mean
    L.orc_sample_texture(C.byref(t), oh.fptr(lut), 0.5, 0.5, oh.fptr(out))
    assert out[0] == pytest.approx(1.5)
    # REPEAT: u = 0 lies halfway between texel 1 (wrapped) and texel 0
    L.orc_sample_texture(C.byref(t), oh.fptr(lut), 0.0, 0.25, oh.fptr(out))
    assert out[0] == pytest.approx(0.5)


def test_srgb_texture_decodes_rgb_not_alpha():
    img = np.array([[[128, 128, 128, 128]]], np.uint8)
    t, keep = _tex(img, abi.RTG_FORMAT_RGBA8_SRGB, abi.RTG_FILTER_NEAREST)
    lut = np.zeros(256, np.float32)
    L.orc_srgb_lut(oh.fptr(lut))
    out = np.zeros(4, np.float32)
    L.orc_sample_texture(C.byref(t), oh.fptr(lut), 0.5, 0.5, oh.fptr(out))
    assert out[0] == lut[128] and out[3] == np.float32(128) / np.float32(255)
