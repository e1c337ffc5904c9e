"""A scene of glTF models (alpha-MASK fence with a nearest-filtered texture,
a normal-mapped torus, a material-less model with the fallback image) traced
on the GPU equals the oracle: hit records and RGBA8 bit for bit, linear RGB
within 1e-4."""
import numpy as np
import pytest
import torch

import oracle_host as oh
from gltf_assets import gltf_scene
from rtg import abi
from rtg.device import DeviceScene, hits_to_numpy

pytestmark = pytest.mark.gpu


def test_gltf_scene_matches_oracle(ctx):
    s = gltf_scene(96, 64)
    ds = DeviceScene(ctx, s)
    hs = oh.HostScene(s)
    for frame in (0, 3):
        rgba, rgb, ph = ds.trace_image(frame_index=frame)
        torch.cuda.synchronize()
        o_rgba, o_rgb, o_hits, cnt = hs.trace_image(frame_index=frame)
        assert cnt.anyhit_calls > 0
        assert np.array_equal(hits_to_numpy(ph).view(np.uint32), o_hits.view(np.uint32))
        assert np.array_equal(rgba.cpu().numpy().view(np.uint32), o_rgba)
        assert np.abs(rgb.cpu().numpy().reshape(-1, 3) - o_rgb).max() <= 1e-4
    ds.destroy()
