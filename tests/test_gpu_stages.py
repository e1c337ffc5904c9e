"""The HIP stages that only some scenes reach, against the oracle on the GPU:
the Portal hit group (closest_hit_portal, shaders/ray-tracing/src/lib.rs:
300-312) in the reference's default-scene layout, and the heatmap mode
(lib.rs:120-124,174-186; heatmap.rs:5-47)."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle_host as oh
from rtg import abi, scenes
from rtg.device import DeviceScene, hits_to_numpy

pytestmark = pytest.mark.gpu

RGB_TOL = 1e-4


def primary_groups(s, hits):
    groups = np.array([ins.hit_group for ins in s.instances])
    inst = hits["instance"]
    return np.where(inst != abi.RTG_NO_HIT, groups[np.minimum(inst, len(groups) - 1)], -1)


def test_portal_scene_matches_oracle(ctx):
    s = scenes.default_scene(width=320, height=180)
    ds = DeviceScene(ctx, s)
    ctx.set_counters(True)
    rgba, rgb, ph = ds.trace_image()
    gc = ctx.counters()
    ctx.set_counters(False)
    torch.cuda.synchronize()
    hs = oh.HostScene(s)
    o_rgba, o_rgb, o_hits, oc = hs.trace_image()
    g_hits = hits_to_numpy(ph)
    assert np.array_equal(g_hits.view(np.uint32), o_hits.view(np.uint32))
    assert np.array_equal(rgba.cpu().numpy().view(np.uint32), o_rgba)
    assert float(np.abs(rgb.cpu().numpy().reshape(-1, 3) - o_rgb).max()) <= RGB_TOL
    # the portal, mirror and textured stages all ran, with their
    # continuation and shadow rays counted alike
    g = primary_groups(s, o_hits)
    assert (g == abi.RTG_HIT_PORTAL).sum() > 500
    assert (g == abi.RTG_HIT_MIRROR).sum() > 500
    assert gc.bounce_rays == oc.bounce_rays and gc.bounce_rays > (g == abi.RTG_HIT_PORTAL).sum()
    assert gc.shadow_rays == oc.shadow_rays and gc.anyhit_calls == oc.anyhit_calls > 0
    ds.destroy()


def heatmap_table():
    """heatmap_temperature(c / 1e6) for every cycle count c in [0, 1e6] (heat
    saturates beyond), from the oracle."""
    c = np.arange(0, 1_000_001, dtype=np.uint32)
    zero = np.zeros((len(c), 3), np.float32)
    out = np.zeros_like(zero)
    oh.lib().orc_heatmap_pixels(c.ctypes.data, zero.ctypes.data, len(c), out.ctypes.data)
    return c, out


def test_heatmap_mode_lies_on_the_ramp(ctx):
    from scipy.spatial import cKDTree
    s = scenes.default_scene(width=160, height=90)
    ds = DeviceScene(ctx, s)
    plain = ds.trace_image()
    opts = s.trace_opts()
    opts.flags = abi.RTG_TRACE_HEATMAP
    heat = ds.trace_image(opts=opts)
    torch.cuda.synchronize()
    # without the flag the frame is the oracle's
    hs = oh.HostScene(s)
    o_rgba, o_rgb, o_hits, _ = hs.trace_image()
    col = plain[1].cpu().numpy().reshape(-1, 3)
    assert np.array_equal(plain[0].cpu().numpy().view(np.uint32), o_rgba)
    assert float(np.abs(col - o_rgb).max()) <= RGB_TOL
    # with it, every pixel is heatmap_temperature(c / 1e6) + colour * 1e-6,
    # bit for bit, for some integer count c of traversal cycles, c > 0
    h_rgb = heat[1].cpu().numpy().reshape(-1, 3)
    c_all, table = heatmap_table()
    k = 48
    _, cand = cKDTree(table.astype(np.float64)).query(h_rgb.astype(np.float64) - col.astype(np.float64) * 1e-6, k=k)
    n = len(h_rgb)
    cyc = c_all[cand.reshape(-1)]
    rep = np.repeat(col, k, axis=0)
    out = np.zeros_like(rep)
    oh.lib().orc_heatmap_pixels(cyc.ctypes.data, rep.ctypes.data, len(cyc), out.ctypes.data)
    match = np.all(out.reshape(n, k, 3).view(np.uint32) == h_rgb.view(np.uint32)[:, None, :], axis=2)
    assert match.any(axis=1).all(), f"{(~match.any(axis=1)).sum()} pixels off the ramp"
    first = np.where(match, cyc.reshape(n, k), np.uint32(0xFFFFFFFF)).min(axis=1)
    assert np.all(first > 0)
    # the RGBA8 output is the sRGB encoding of that colour
    enc = np.zeros(3, np.float32)
    got = heat[0].cpu().numpy().view(np.uint32)
    for i in range(0, n, 7):
        oh.lib().orc_linear_to_srgb(oh.fptr(np.ascontiguousarray(h_rgb[i])), oh.fptr(enc))
        q = (np.clip(enc, 0.0, 1.0) * np.float32(255.0) + np.float32(0.5)).astype(np.uint32)
        assert got[i] == q[0] | (q[1] << 8) | (q[2] << 16) | (255 << 24)
    # the uniforms' show_heatmap switches the same mode
    u = s.uniforms()
    u.show_heatmap = 1
    out2 = ds.make_outputs()
    j = ds.job(out2, uniforms=u)
    from rtg.lib import check, load
    check(load().rtg_trace_image(ctx.ptr, C.byref(j), None), "rtg_trace_image")
    torch.cuda.synchronize()
    assert not np.array_equal(out2[1].cpu().numpy(), plain[1].cpu().numpy())
    ds.destroy()


def odd_texture_scene():
    """The small scene with textures whose sizes are not powers of two: the
    ground's base colour a 24x20 sRGB texture filtered bilinearly (its
    UVs run to 4.3, and bilinear taps at the borders index -1), its
    metallic-roughness a 17x13 UNORM texture sampled nearest, the blue-noise
    mask 48x40, with two soft-shadow rays per hit."""
    s = scenes.small_scene(width=160, height=120, n_tori=4)
    rng = np.random.default_rng(0x0DD)
    base = rng.integers(0, 256, (20, 24, 4), dtype=np.uint8)
    mr = rng.integers(0, 256, (13, 17, 4), dtype=np.uint8)
    # every 8-bit code appears in the UNORM textures' channels
    mr.reshape(-1)[:256] = np.arange(256, dtype=np.uint8)
    bn = rng.integers(0, 256, (40, 48, 4), dtype=np.uint8)
    bn.reshape(-1)[:256] = np.arange(256, dtype=np.uint8)
    d = s.add_texture(scenes.TextureData(24, 20, abi.RTG_FORMAT_RGBA8_SRGB, abi.RTG_FILTER_LINEAR, base, "odd base"))
    m = s.add_texture(scenes.TextureData(17, 13, abi.RTG_FORMAT_RGBA8_UNORM, abi.RTG_FILTER_NEAREST, mr, "odd mr"))
    s.textures[2] = scenes.TextureData(48, 40, abi.RTG_FORMAT_RGBA8_UNORM, abi.RTG_FILTER_NEAREST, bn, "odd noise")
    g_mesh = s.add_mesh(scenes.quad(40.0, 40.0, 4.3, "odd ground"))
    s.instances[0] = scenes.InstanceData(scenes.affine((0.0, 0.0, 0.0)), s.add_model(g_mesh, [(d, m, -1)]), g_mesh)
    s.sun_radius = float(np.tan(np.radians(1.0)))
    s.shadow_rays = 2
    return s


def test_textures_of_any_size_match_oracle(ctx):
    """Texel wrapping and 8-bit decoding for texture sizes that are not
    powers of two, with negative bilinear taps, against the oracle."""
    s = odd_texture_scene()
    ds = DeviceScene(ctx, s)
    rgba, rgb, ph = ds.trace_image()
    torch.cuda.synchronize()
    o_rgba, o_rgb, o_hits, _ = oh.HostScene(s).trace_image()
    assert np.array_equal(hits_to_numpy(ph).view(np.uint32), o_hits.view(np.uint32))
    assert np.array_equal(rgba.cpu().numpy().view(np.uint32), o_rgba)
    assert float(np.abs(rgb.cpu().numpy().reshape(-1, 3) - o_rgb).max()) <= RGB_TOL
    # the ground fills much of the frame
    assert (o_hits["instance"] == 0).mean() > 0.3
    ds.destroy()
