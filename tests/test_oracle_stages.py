"""Pins the oracle's stage functions, one by one, against what the geometry
and the mathematics fix: closed forms evaluated in double precision, exact
known answers and invariants. The oracle's per-pixel trace calls these same
functions, and the GPU path equals the oracle bit for bit on whole frames, so
each pin holds for the HIP stages too.

Stages (reference file:line):
  camera ray          shaders/ray-tracing/src/lib.rs:126-142
  primary miss        lib.rs:38-51
  terminator fix      shaders/closest_hit_textured.glsl:13-39
  blue-noise numbers  closest_hit_textured.glsl:96-120
  shading normal      closest_hit_textured.glsl:65-72,123-157
  textured colour     closest_hit_textured.glsl:203-225 (with pbr.glsl)
  mirror              shaders/closest_hit_mirror.glsl:11-29
  portal              lib.rs:300-312
  heatmap             shaders/ray-tracing/src/heatmap.rs:5-47, lib.rs:174-186
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_host as oh
from rtg import abi, scenes
from test_oracle_math import brdf_double

L = oh.lib()


def f32a(*xs):
    return np.ascontiguousarray(np.array(xs, dtype=np.float32).reshape(-1))


def out3():
    return np.zeros(3, np.float32)


def uniforms(view_inverse, proj_inverse):
    """rtg_uniforms with two 4x4 matrices given in mathematical (row, column)
    order; the struct stores them column-major."""
    U = abi.Uniforms()
    U.view_inverse[:] = np.asarray(view_inverse, np.float32).T.reshape(-1).tolist()
    U.proj_inverse[:] = np.asarray(proj_inverse, np.float32).T.reshape(-1).tolist()
    return U


def camera_ray(U, w, h, px, py, jx=0.5, jy=0.5):
    o, d = out3(), out3()
    L.orc_camera_ray(C.byref(U), w, h, px, py, jx, jy, oh.fptr(o), oh.fptr(d))
    return o, d


def rot(axis, angle):
    a = np.asarray(axis, np.float64)
    a /= np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def o2w_of(R, T=(0.0, 0.0, 0.0)):
    """Row-major 3x4 object-to-world matrix (float32) of x -> R x + T."""
    m = np.zeros((3, 4), np.float64)
    m[:, :3] = R
    m[:, 3] = T
    return np.ascontiguousarray(m.astype(np.float32).reshape(-1))


# ---------------------------------------------------------------------------
# camera ray
# ---------------------------------------------------------------------------
# proj_inverse with P^-1 (x, y, 1, 1) = (a x, b y, -1, c + d): the view-space
# direction through render coordinates (x, y) is (a x, b y, -1); c + d < 0
# makes a perspective divide flip the direction.
A_, B_ = 0.8, 0.45
PINV = [[A_, 0, 0, 0], [0, B_, 0, 0], [0, 0, 0, -1], [0, 0, 0.5, -3.0]]


def expected_view_dir(w, h, px, py, jx=0.5, jy=0.5):
    x = 2.0 * (px + jx) / w - 1.0
    y = 2.0 * (py + jy) / h - 1.0
    v = np.array([A_ * x, B_ * y, -1.0])
    return v / np.linalg.norm(v)


def test_camera_ray_identity_view_is_the_normalised_projected_pixel_centre():
    U = uniforms(np.eye(4), PINV)
    w, h = 64, 48
    for py in range(0, h, 5):
        for px in range(0, w, 3):
            o, d = camera_ray(U, w, h, px, py)
            assert np.all(o == 0.0)
            np.testing.assert_allclose(d, expected_view_dir(w, h, px, py), rtol=0, atol=5e-7)
    # pixel (0, 0) looks towards -x, -y: render coordinates grow with x and y
    _, d = camera_ray(U, w, h, 0, 0)
    assert d[0] < 0 and d[1] < 0
    # jittered sample point
    _, d = camera_ray(U, w, h, 10, 20, 0.25, 0.75)
    np.testing.assert_allclose(d, expected_view_dir(w, h, 10, 20, 0.25, 0.75), atol=5e-7)


def test_camera_ray_view_transform_moves_origin_and_rotates_direction():
    R = rot((0.3, 1.0, -0.2), 2.1)
    T = np.array([1.5, 2.0, -5.0])
    V = np.eye(4)
    V[:3, :3] = R
    V[:3, 3] = T
    U = uniforms(V, PINV)
    Rf = R.astype(np.float32).astype(np.float64)
    for px, py in ((0, 0), (31, 7), (63, 47), (12, 40)):
        o, d = camera_ray(U, 64, 48, px, py)
        assert np.array_equal(o, T.astype(np.float32))
        np.testing.assert_allclose(d, Rf @ expected_view_dir(64, 48, px, py), atol=1e-6)
        assert abs(np.linalg.norm(d.astype(np.float64)) - 1.0) < 1e-6


def test_camera_ray_half_angles_of_the_scene_projection():
    # the scenes' camera: 60 degree vertical field of view; a ray through
    # render coordinates (x, y) makes tan(angle) = |y| tan(fov/2) with the
    # optical axis vertically and |x| aspect tan(fov/2) horizontally
    fov, w, h = math.radians(60.0), 40, 30
    U = uniforms(np.eye(4), scenes.proj_inverse_reversed_z(fov, w / h))
    _, axis = camera_ray(U, 3, 3, 1, 1)  # the centre of a 3x3 launch: (x, y) = (0, 0)
    axis = axis.astype(np.float64)
    for px, py in ((0, 0), (39, 29), (5, 17)):
        _, d = camera_ray(U, w, h, px, py)
        d = d.astype(np.float64)
        x = 2.0 * (px + 0.5) / w - 1.0
        y = 2.0 * (py + 0.5) / h - 1.0
        along = d @ axis
        side = d - along * axis
        assert abs(side[1]) / along == pytest.approx(abs(y) * math.tan(fov / 2), rel=1e-5)
        assert math.hypot(side[0], side[2]) / along == pytest.approx(abs(x) * (w / h) * math.tan(fov / 2), rel=1e-5)


# ---------------------------------------------------------------------------
# primary miss
# ---------------------------------------------------------------------------
def test_primary_miss_sun_disk_threshold():
    sun = np.array([0.3, 0.8, 0.5])
    sun /= np.linalg.norm(sun)
    radius = 0.05
    perp = np.cross(sun, [0.0, 0.0, 1.0])
    perp /= np.linalg.norm(perp)
    c = out3()
    for angle, white in ((0.0, True), (0.99 * radius, True), (1.01 * radius, False), (0.5, False),
                         (math.pi, False)):
        d = math.cos(angle) * sun + math.sin(angle) * perp
        L.orc_primary_miss(oh.fptr(f32a(*d)), oh.fptr(f32a(*sun)), radius, oh.fptr(c))
        want = np.array([1, 1, 1], np.float32) if white else np.array([0, 0, 0.05], np.float32)
        assert np.array_equal(c, want), angle


# ---------------------------------------------------------------------------
# shadow-terminator fix
# ---------------------------------------------------------------------------
def terminator(P, N, bc):
    o = out3()
    L.orc_terminator_fix(oh.fptr(f32a(*np.ravel(P))), oh.fptr(f32a(*np.ravel(N))), oh.fptr(f32a(*bc)), oh.fptr(o))
    return o.astype(np.float64)


def cap_triangle(alpha):
    """Three points of the unit sphere at polar angle alpha around +y, 120
    degrees apart."""
    return np.array([[math.sin(alpha) * math.cos(p), math.cos(alpha), math.sin(alpha) * math.sin(p)]
                     for p in (0.0, 2 * math.pi / 3, 4 * math.pi / 3)])


def test_terminator_fix_flat_triangle_keeps_the_hit_point():
    rng = np.random.default_rng(5)
    for _ in range(50):
        P = rng.uniform(-2, 2, (3, 3)).astype(np.float32).astype(np.float64)
        n = np.cross(P[1] - P[0], P[2] - P[0])
        n /= np.linalg.norm(n)
        bc = rng.dirichlet([1, 1, 1])
        bc32 = bc.astype(np.float32).astype(np.float64)
        pos = bc32 @ P
        np.testing.assert_allclose(terminator(P, [n, n, n], bc32), pos, atol=2e-6)


def test_terminator_fix_sphere_cap_centroid_closed_form():
    # vertex normals = vertex positions on the unit sphere; at the centroid c
    # every offset is c - P_i - d P_i with d = c.P_i - 1 = (2k - 2)/3,
    # k = P_i.P_j, so the fixed point is c (5 - 2k) / 3
    for alpha in (0.05, 0.2, 0.5):
        P = cap_triangle(alpha).astype(np.float32).astype(np.float64)
        k = float(P[0] @ P[1])
        c = P.mean(axis=0)
        got = terminator(P, P, [1 / 3, 1 / 3, 1 / 3])
        np.testing.assert_allclose(got, c * (5 - 2 * k) / 3, atol=2e-6)
        assert np.linalg.norm(got) > np.linalg.norm(c)


def test_terminator_fix_at_a_vertex_and_on_a_concave_patch():
    P = cap_triangle(0.3).astype(np.float32).astype(np.float64)
    for k in range(3):
        bc = [0.0, 0.0, 0.0]
        bc[k] = 1.0
        np.testing.assert_allclose(terminator(P, P, bc), P[k], atol=1e-6)
    # normals pointing to the sphere's centre: every (P - P_i).n_i >= 0, no offset
    rng = np.random.default_rng(6)
    for _ in range(20):
        bc = rng.dirichlet([1, 1, 1]).astype(np.float32).astype(np.float64)
        np.testing.assert_allclose(terminator(P, -P, bc), bc @ P, atol=1e-6)


# ---------------------------------------------------------------------------
# blue-noise numbers
# ---------------------------------------------------------------------------
def coord_texture():
    """64x64 RGBA32F texture whose red channel is (x + 64 y) / 4096 (exact)."""
    img = np.zeros((64, 64, 4), np.float32)
    yy, xx = np.mgrid[0:64, 0:64]
    img[..., 0] = (xx + 64 * yy) / 4096.0
    t = abi.Texture()
    t.data = img.ctypes.data
    t.width = t.height = 64
    t.format = abi.RTG_FORMAT_RGBA32_FLOAT
    t.filter = abi.RTG_FILTER_NEAREST
    return t, img


def blue_noise(t, px, py, i, frame):
    out = np.zeros(2, np.float32)
    L.orc_blue_noise(C.byref(t), oh.fptr(np.zeros(256, np.float32)), px, py, i, frame, oh.fptr(out))
    return out


def test_blue_noise_texel_offsets():
    t, keep = coord_texture()
    for px, py in ((0, 0), (5, 9), (63, 63), (130, 77), (1919, 1079)):
        for i in range(6):
            r = blue_noise(t, px, py, i, 0)
            got = [(int(round(v * 4096)) % 64, int(round(v * 4096)) // 64) for v in r]
            assert got[0] == ((px + 2 * i * 13) % 64, (py + 2 * i * 41) % 64)
            assert got[1] == ((px + (2 * i + 1) * 13) % 64, (py + (2 * i + 1) * 41) % 64)


def test_blue_noise_golden_ratio_animation():
    t, keep = coord_texture()
    # texel (0, 0) holds 0: frame f gives fract(f * 0.618033988749)
    assert blue_noise(t, 0, 0, 0, 1)[0] == np.float32(0.618033988749)
    assert blue_noise(t, 0, 0, 0, 2)[0] == pytest.approx(2 * 0.618033988749 - 1, abs=2e-7)
    assert blue_noise(t, 0, 0, 0, 3)[0] == pytest.approx(3 * 0.618033988749 - 1, abs=3e-7)
    # period 32 in the frame index
    for px, py, i in ((3, 4, 0), (40, 1, 2), (17, 63, 5)):
        for f in (0, 1, 7, 31):
            assert np.array_equal(blue_noise(t, px, py, i, f), blue_noise(t, px, py, i, f + 32))
            assert np.array_equal(blue_noise(t, px, py, i, f), blue_noise(t, px, py, i, f + 64))
        assert not np.array_equal(blue_noise(t, px, py, i, 1), blue_noise(t, px, py, i, 2))


# ---------------------------------------------------------------------------
# shading normal
# ---------------------------------------------------------------------------
def shading_normal(o2w, nrm, P, UV, map_rgb=None):
    o = out3()
    mp = oh.fptr(f32a(*map_rgb)) if map_rgb is not None else None
    keep = f32a(*map_rgb) if map_rgb is not None else None
    if keep is not None:
        mp = oh.fptr(keep)
    L.orc_shading_normal(oh.fptr(o2w), oh.fptr(f32a(*nrm)), oh.fptr(f32a(*np.ravel(P))),
                         oh.fptr(f32a(*np.ravel(UV))), mp, oh.fptr(o))
    return o


def test_shading_normal_uses_the_inverse_transpose():
    rng = np.random.default_rng(8)
    P = np.zeros((3, 3))
    UV = np.zeros((3, 2))
    for _ in range(50):
        R = rot(rng.normal(size=3), rng.uniform(0, 6))
        S = np.diag(rng.uniform(0.3, 3.0, 3))
        M = R @ S
        o2w = o2w_of(M, rng.uniform(-5, 5, 3))
        Mf = o2w.reshape(3, 4)[:, :3].astype(np.float64)
        n = rng.normal(size=3)
        want = np.linalg.inv(Mf).T @ n
        want /= np.linalg.norm(want)
        np.testing.assert_allclose(shading_normal(o2w, n, P, UV), want, atol=2e-6)


def tangent_triangle(U, V, a, b, uv0, duv1, duv2):
    """Triangle whose positions are Q + U a u + V b v of its uvs (u, v)."""
    UV = np.array([uv0, np.add(uv0, duv1), np.add(uv0, duv2)], np.float64)
    Q = np.array([0.3, -0.2, 0.7])
    P = np.array([Q + U * a * (uv[0] - uv0[0]) + V * b * (uv[1] - uv0[1]) for uv in UV])
    return P, UV


def test_shading_normal_flat_map_texel_is_the_vertex_normal():
    U = np.array([1.0, 0.0, 0.0])
    V = np.array([0.0, 0.0, -1.0])
    N = np.cross(U, V)
    P, UV = tangent_triangle(U, V, 1.0, 1.0, (0.1, 0.2), (0.5, 0.0), (0.0, 0.5))
    o2w = o2w_of(rot((1, 2, 3), 0.7), (1, 2, 3))
    plain = shading_normal(o2w, N, P, UV)
    assert np.array_equal(shading_normal(o2w, N, P, UV, (0.5, 0.5, 1.0)), plain)


@pytest.mark.parametrize("a,b", [(1.0, 1.0), (2.0, 0.5), (0.25, 1.5)])
def test_shading_normal_map_follows_the_uv_tangent_frame(a, b):
    # positions Q + U a u + V b v: the cotangent frame is T = b det U,
    # B = a det V (det > 0 for counter-clockwise uvs), scaled so that the
    # longer of the two has unit length; the map vector m = normalize(rgb*2-1)
    # gives the model-space normal T' m.x + B' m.y + N m.z
    rng = np.random.default_rng(9)
    for _ in range(30):
        R0 = rot(rng.normal(size=3), rng.uniform(0, 6))
        U, V = R0[:, 0], R0[:, 1]
        N = np.cross(U, V)
        P, UV = tangent_triangle(U, V, a, b, rng.uniform(0, 1, 2), (0.3, 0.05), (0.02, 0.4))
        rgb = rng.uniform(0.2, 0.8, 3)
        rgb[2] = rng.uniform(0.7, 1.0)
        m = rgb.astype(np.float32).astype(np.float64) * 2 - 1
        m /= np.linalg.norm(m)
        s = max(a, b)
        local = U * (b / s) * m[0] + V * (a / s) * m[1] + N * m[2]
        R = rot(rng.normal(size=3), rng.uniform(0, 6))
        o2w = o2w_of(R, rng.uniform(-3, 3, 3))
        Rf = o2w.reshape(3, 4)[:, :3].astype(np.float64)
        want = np.linalg.inv(Rf).T @ local
        want /= np.linalg.norm(want)
        got = shading_normal(o2w, N.astype(np.float32), P.astype(np.float32), UV.astype(np.float32), rgb)
        np.testing.assert_allclose(got, want, atol=1e-5)


# ---------------------------------------------------------------------------
# textured colour: BRDF + ambient
# ---------------------------------------------------------------------------
def textured_colour(n, ray_dir, sun, rough, base, metallic, sf):
    o = out3()
    L.orc_textured_colour(oh.fptr(f32a(*n)), oh.fptr(f32a(*ray_dir)), oh.fptr(f32a(*sun)), rough,
                          oh.fptr(f32a(*base)), metallic, sf, oh.fptr(o))
    return o


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_textured_colour_matches_filament_plus_ambient_in_double():
    rng = np.random.default_rng(10)
    for _ in range(300):
        n = unit(rng.normal(size=3))
        view = unit(rng.normal(size=3))
        if n @ view < 0:
            view = -view
        sun = unit(rng.normal(size=3))
        rough, metallic, sf = rng.uniform(0.2, 1.0), rng.uniform(), rng.choice([0.25, 0.5, 0.75, 1.0])
        base = rng.uniform(0, 1, 3)
        n32, v32, s32, b32 = (np.asarray(x, np.float32) for x in (n, view, sun, base))
        got = textured_colour(n32, -v32, s32, np.float32(rough), b32, np.float32(metallic), sf)
        want = brdf_double(n32.astype(np.float64), v32.astype(np.float64), s32.astype(np.float64),
                           float(np.float32(rough)), b32.astype(np.float64), float(np.float32(metallic)), 0.5, sf)
        want = want + 0.1 * b32.astype(np.float64)
        np.testing.assert_allclose(got, want, rtol=2e-4, atol=2e-6)


def test_textured_colour_in_full_shadow_is_the_ambient_term():
    rng = np.random.default_rng(11)
    for _ in range(50):
        n = unit(rng.normal(size=3))
        base = rng.uniform(0, 1, 3).astype(np.float32)
        got = textured_colour(n, -n, unit(rng.normal(size=3)), 0.5, base, 0.3, 0.0)
        assert np.array_equal(got, np.float32(0.1) * base)
        # the direct term is linear in the mean visibility
        full = textured_colour(n, -n, unit(n + rng.normal(size=3) * 0.3), 0.5, base, 0.3, 1.0)
        quarter = textured_colour(n, -n, unit(n + rng.normal(size=3) * 0.3), 0.5, base, 0.3, 0.25)
        assert np.all(full >= np.float32(0.1) * base) and np.all(quarter >= np.float32(0.1) * base)


# ---------------------------------------------------------------------------
# mirror and portal continuations
# ---------------------------------------------------------------------------
def mirror(o2w, n, o, d, t):
    no, nd = out3(), out3()
    L.orc_mirror(oh.fptr(o2w), oh.fptr(f32a(*n)), oh.fptr(f32a(*o)), oh.fptr(f32a(*d)), t, oh.fptr(no), oh.fptr(nd))
    return no, nd


def test_mirror_reflects_about_the_plane_exactly():
    o2w = o2w_of(np.eye(3))
    d = f32a(*unit([1.0, -1.0, -1.0]))
    o = f32a(0.5, 1.0, 2.0)
    no, nd = mirror(o2w, (0.0, 0.0, 1.0), o, d, 2.5)
    assert np.array_equal(nd, np.array([d[0], d[1], -d[2]], np.float32))
    assert np.array_equal(no, o + d * np.float32(2.5))


def test_mirror_preserves_length_and_angle_under_instance_transforms():
    rng = np.random.default_rng(12)
    for _ in range(100):
        R = rot(rng.normal(size=3), rng.uniform(0, 6))
        S = np.diag(rng.uniform(0.5, 2.0, 3))
        o2w = o2w_of(R @ S, rng.uniform(-4, 4, 3))
        Mf = o2w.reshape(3, 4)[:, :3].astype(np.float64)
        n = rng.normal(size=3) * rng.uniform(0.5, 2.0)  # an unnormalised interpolated normal
        nw = unit(np.linalg.inv(Mf).T @ n)
        d = f32a(*unit(rng.normal(size=3)))
        o = f32a(*rng.uniform(-3, 3, 3))
        t = np.float32(rng.uniform(0.1, 20.0))
        no, nd = mirror(o2w, n.astype(np.float32), o, d, t)
        d64, nd64 = d.astype(np.float64), nd.astype(np.float64)
        assert abs(np.linalg.norm(nd64) - 1.0) < 2e-6
        assert nd64 @ nw == pytest.approx(-(d64 @ nw), abs=2e-6)
        np.testing.assert_allclose(np.cross(nd64 - d64, nw), 0.0, atol=4e-6)  # d' - d along the normal
        np.testing.assert_allclose(no, o.astype(np.float64) + d64 * float(t), rtol=1e-6, atol=1e-5)


def test_portal_moves_the_hit_point_up_five():
    no, nd = out3(), out3()
    for o, d, t, want in (((0.5, 1.25, -2.0), (0.0, -1.0, 0.0), 2.5, (0.5, 3.75, -2.0)),
                          ((1.0, 2.0, 3.0), (0.0, 0.0, 1.0), 2.0, (1.0, 7.0, 5.0)),
                          ((-4.0, 1.0, 0.25), (1.0, 0.0, 0.0), 0.5, (-3.5, 6.0, 0.25))):
        L.orc_portal(oh.fptr(f32a(*o)), oh.fptr(f32a(*d)), t, oh.fptr(no), oh.fptr(nd))
        assert np.array_equal(no, f32a(*want))
        assert np.array_equal(nd, f32a(*d))


# ---------------------------------------------------------------------------
# heatmap ramp (heatmap.rs:5-47) at exact points: heat * 10 = h; weights
# wc = s(hf-0.8, hf+0.8, h) (1 - s(hc-0.8, hc+0.8, h)), wp = 1 - s(hf-0.8, ...),
# wn = s(hc-0.8, ...), s = smoothstep. At integer h: s = 1/2 on both sides, so
# (wc, wp, wn) = (1/4, 1/2, 1/2). At h = k + 1/2: t = 13/16 and 3/16, so
# s = 1859/2048 and 189/2048.
# ---------------------------------------------------------------------------
RAMP = np.array([[0, 2, 91], [0, 108, 251], [0, 221, 221], [51, 221, 0], [255, 252, 0], [255, 180, 0],
                 [255, 104, 0], [226, 22, 0], [191, 0, 83], [145, 0, 65]], np.float64) / 255.0
S_HI, S_LO = 1859 / 2048, 189 / 2048


def heatmap(heat):
    o = out3()
    L.orc_heatmap_temperature(heat, oh.fptr(o))
    return o


@pytest.mark.parametrize("heat,cur,prv,nxt,w", [
    (0.0, 0, 0, 1, (0.25, 0.5, 0.5)),
    (-3.0, 0, 0, 1, (0.25, 0.5, 0.5)),
    (0.5, 5, 4, 6, (0.25, 0.5, 0.5)),
    (0.25, 2, 1, 3, (S_HI * S_HI, S_LO, S_LO)),
    (0.75, 7, 6, 8, (S_HI * S_HI, S_LO, S_LO)),
    (0.05, 0, 0, 1, (S_HI * S_HI, S_LO, S_LO)),
    (0.95, 9, 8, 9, (S_HI * S_HI, S_LO, S_LO)),
    # heat 1 makes h = 10: the reference indexes one past its 10 colours
    # there; the oracle (and the device) take the last colour
    (1.0, 9, 9, 9, (0.25, 0.5, 0.5)),
    (7.0, 9, 9, 9, (0.25, 0.5, 0.5)),
])
def test_heatmap_ramp_exact_points(heat, cur, prv, nxt, w):
    want = np.clip(w[0] * RAMP[cur] + w[1] * RAMP[prv] + w[2] * RAMP[nxt], 0.0, 1.0)
    np.testing.assert_allclose(heatmap(heat), want, rtol=0, atol=1e-6)


def test_heatmap_pixels_add_a_millionth_of_the_colour():
    cycles = np.array([0, 500_000, 250_000, 1_000_000, 4_000_000_000], np.uint32)
    rgb = np.array([[0.5, 0.25, 1.0], [1.0, 1.0, 1.0], [0.0, 2.0, 0.0], [3.0, 0.0, 0.5], [0.2, 0.4, 0.6]],
                   np.float32)
    out = np.zeros_like(rgb)
    L.orc_heatmap_pixels(cycles.ctypes.data, rgb.ctypes.data, len(cycles), out.ctypes.data)
    for c, col, o in zip(cycles, rgb, out):
        hm = heatmap(np.float32(c) / np.float32(1e6))
        assert np.array_equal(o, hm + col * np.float32(1e-6))
    np.testing.assert_allclose(out[1], np.clip(0.25 * RAMP[5] + 0.5 * RAMP[4] + 0.5 * RAMP[6], 0, 1) + 1e-6,
                               atol=1e-6)
