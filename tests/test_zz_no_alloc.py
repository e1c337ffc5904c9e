"""rtg_trace_rays_static (the ray query whose scratch is a static device
array) against rtg_trace_rays: hit records and occlusion flags bit for bit,
on two calls in a row, the second with fewer rays than the first."""
import ctypes as C

import numpy as np
import pytest
import torch

from rtg import abi, scenes
from rtg.device import DeviceScene, hits_to_numpy, rays_to_device
from rtg.lib import check, load

pytestmark = pytest.mark.gpu


def query(entry, ds, rays_dev, n, mode):
    if mode == abi.RTG_QUERY_CLOSEST:
        out = torch.full((n * 6,), -1, dtype=torch.int32, device=ds.device)
    else:
        out = torch.full((n,), 7, dtype=torch.int32, device=ds.device)
    hits = out.data_ptr() if mode == abi.RTG_QUERY_CLOSEST else 0
    occ = out.data_ptr() if mode == abi.RTG_QUERY_OCCLUSION else 0
    check(entry(ds.ctx.ptr, ds.tlas, C.byref(ds.bindings), rays_dev.data_ptr(), n, mode, 0, 0xFF, 0, hits, occ,
                None), "ray query")
    torch.cuda.synchronize()
    a = out.cpu().numpy()
    return hits_to_numpy(out).view(np.uint32) if mode == abi.RTG_QUERY_CLOSEST else a.astype(np.uint32)


@pytest.mark.parametrize("scene_kind", ["soup", "fences"])
def test_static_query_equals_allocating_query(ctx, scene_kind):
    if scene_kind == "soup":
        s = scenes.incoherent_scene(n_tris=200_000)
        lo, hi = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    else:
        # instanced objects, alpha-clipped fences (any-hit) and mirrors
        s = scenes.large_scene(width=64, height=48, n_objects=96, n_fences=12, n_mirrors=6, n_soup_meshes=4,
                               extent=10.0)
        lo, hi = (-10.0, 0.0, -10.0), (10.0, 3.0, 10.0)
    ds = DeviceScene(ctx, s)
    L = load()
    # the first call is larger than the static grid (512 workgroups of 256
    # lanes), the second smaller
    for n, seed in ((300_000, 11), (40_000, 12)):
        print(f"static query: scene {scene_kind}, {n} rays, seed {seed}", flush=True)
        rays = scenes.incoherent_rays(n, lo, hi, seed=seed)
        rd = rays_to_device(rays)
        for mode in (abi.RTG_QUERY_CLOSEST, abi.RTG_QUERY_OCCLUSION):
            ref = query(L.rtg_trace_rays, ds, rd, n, mode)
            got = query(L.rtg_trace_rays_static, ds, rd, n, mode)
            assert np.array_equal(got, ref), (scene_kind, n, mode)
            if mode == abi.RTG_QUERY_CLOSEST:
                hit = ref.reshape(n, 6)[:, 1] != abi.RTG_NO_HIT
                assert 0 < hit.sum() < n
            else:
                assert 0 < ref.sum() < n
    ds.destroy()


def test_static_query_rejects_bad_arguments(ctx):
    s = scenes.incoherent_scene(n_tris=2_000)
    ds = DeviceScene(ctx, s)
    L = load()
    # not a TLAS of this context, and an unknown mode: refused on the host
    st = L.rtg_trace_rays_static(ctx.ptr, 12345, None, 0, 1, 0, 0, 0xFF, 0, 0, 0, None)
    assert st == abi.RTG_ERROR_NOT_FOUND
    st = L.rtg_trace_rays_static(ctx.ptr, ds.tlas, None, 0, 0, 7, 0, 0xFF, 0, 0, 0, None)
    assert st == abi.RTG_ERROR_INVALID_ARGUMENT
    ds.destroy()
