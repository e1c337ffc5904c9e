"""Request serving over resident scenes (rtg_server_* of include/rtg.h).

The server keeps every registered scene's BLASes, TLAS, textures and tables
resident in HBM; each request names a scene, an image size, options, its own
instance transforms (refit of a private TLAS copy) and its output buffers.
A native worker batches pending requests continuously.
"""
import ctypes as C

import numpy as np

from . import abi
from .lib import check, load


class Server:
    def __init__(self, ctx, budget_bytes=8 << 30, max_batch=16, max_inflight=2, max_samples=0):
        cfg = abi.ServerConfig(int(budget_bytes), int(max_batch), int(max_inflight), int(max_samples))
        self.ptr = C.c_void_p()
        check(load().rtg_server_create(ctx.ptr, C.byref(cfg), C.byref(self.ptr)), "rtg_server_create")
        self.scenes = []

    def add_scene(self, device_scene):
        sid = C.c_uint32()
        check(load().rtg_server_add_scene(self.ptr, device_scene.tlas, C.byref(device_scene.bindings),
                                          C.byref(sid)), "rtg_server_add_scene")
        self.scenes.append(device_scene)
        return sid.value

    def submit(self, req):
        t = C.c_uint64()
        st = load().rtg_server_submit(self.ptr, C.byref(req), C.byref(t))
        return st, t.value

    def cancel(self, ticket):
        """RTG_OK if the request was still pending (it completes with
        RTG_ERROR_CANCELLED), else RTG_ERROR_NOT_FOUND."""
        return load().rtg_server_cancel(self.ptr, int(ticket))

    def clock(self):
        return float(load().rtg_server_clock(self.ptr))

    def poll(self, wait_all=False, capacity=4096):
        out = (abi.Completion * capacity)()
        n = C.c_uint32()
        check(load().rtg_server_poll(self.ptr, out, capacity, C.byref(n), 1 if wait_all else 0), "rtg_server_poll")
        return [out[i] for i in range(n.value)]

    def replay(self, requests, arrivals):
        n = len(requests)
        reqs = (abi.Request * n)(*requests)
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        out = (abi.Completion * n)()
        check(load().rtg_server_replay(self.ptr, reqs, arr.ctypes.data_as(C.POINTER(C.c_double)), n, out),
              "rtg_server_replay")
        return [out[i] for i in range(n)]

    def stats(self):
        s = abi.ServerStats()
        check(load().rtg_server_get_stats(self.ptr, C.byref(s)), "rtg_server_get_stats")
        return s.as_dict()

    def close(self):
        if self.ptr:
            load().rtg_server_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_request(scene_id, device_scene, width, height, spp, shadow_rays, max_traces, frame, instances=None,
                 outputs=None, tag=0, deadline_s=0.0):
    """A Request with the scene's camera and sun; `instances` is a device
    instance buffer (or None to trace the scene as built)."""
    r = abi.Request()
    r.scene = scene_id
    r.width, r.height = width, height
    r.opts.spp, r.opts.shadow_rays, r.opts.max_traces, r.opts.flags = spp, shadow_rays, max_traces, 0
    s = device_scene.scene
    from .scenes import proj_inverse_reversed_z
    import math
    r.uniforms = abi.uniforms_from(s.view_inverse, proj_inverse_reversed_z(math.radians(60.0), width / height),
                                   s.sun_dir, s.sun_radius, s.blue_noise_index, s.lut_index, frame)
    r.instances = instances.data_ptr() if instances is not None else 0
    if outputs is not None:
        rgba, rgb, hits = outputs
        r.out_rgba8 = rgba.data_ptr() if rgba is not None else 0
        r.out_rgb = rgb.data_ptr() if rgb is not None else 0
        r.out_primary_hits = hits.data_ptr() if hits is not None else 0
    r.tag = tag
    r.deadline_s = float(deadline_s)
    return r
