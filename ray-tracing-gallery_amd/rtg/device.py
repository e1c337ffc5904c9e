"""Device-resident scenes for the HIP trace step.

DeviceScene uploads a scenes.Scene to HBM (torch tensors as plain device
memory), builds one BLAS per mesh (≙ Model::new -> build_blas,
src/util_structs.rs:1158-1236), the ModelInfo/GeometryInfo tables
(src/util_structs.rs:1137-1143, 1180-1197), the texture table
(≙ ImageManager, src/util_structs.rs:1296-1377), the instance buffer and the
TLAS (≙ build_tlas, src/main.rs:510-542), and traces frames through the C ABI
(≙ PerFrameResources::record -> vkCmdTraceRaysKHR).
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from .lib import Context, check, load


def _dev(arr, device):
    t = torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).reshape(-1).copy())
    return t.to(device)


def _stream_ptr(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


class DeviceScene:
    def __init__(self, ctx: Context, scene, device="cuda:0", stream=None):
        L = load()
        self.ctx = ctx
        self.scene = scene
        self.device = torch.device(device)
        self.keep = []
        sp = _stream_ptr(stream)
        # meshes -> vertex streams, index buffers, BLASes
        self.mesh_bufs = []
        self.blas = []
        for m in scene.meshes:
            pos = _dev(m.positions, self.device)
            nrm = _dev(m.normals, self.device)
            uv = _dev(m.uvs, self.device)
            idx = [_dev(np.ascontiguousarray(ix, dtype=np.uint32), self.device) for ix, _ in m.geometries]
            self.mesh_bufs.append((pos, nrm, uv, idx))
            geoms = (abi.Geometry * max(1, len(m.geometries)))()
            for g, (ix, op) in enumerate(m.geometries):
                geoms[g].vertex_data = pos.data_ptr()
                geoms[g].vertex_stride = 12
                geoms[g].max_vertex = max(0, len(m.positions) - 1)
                geoms[g].index_data = idx[g].data_ptr()
                geoms[g].index_count = ix.size
                geoms[g].flags = abi.RTG_GEOMETRY_OPAQUE if op else 0
            out = C.c_uint64()
            check(L.rtg_blas_build(ctx.ptr, C.cast(geoms, C.c_void_p), len(m.geometries),
                                   abi.RTG_BUILD_PREFER_FAST_TRACE, C.byref(out), sp), "rtg_blas_build")
            self.blas.append(out.value)
        # textures
        tex = np.zeros(max(1, len(scene.textures)), dtype=abi.TEXTURE_DTYPE)
        self.tex_bufs = []
        for k, t in enumerate(scene.textures):
            buf = _dev(t.data, self.device)
            self.tex_bufs.append(buf)
            tex[k] = (buf.data_ptr(), t.width, t.height, t.format, t.filter)
        self.tex_table = _dev(tex, self.device)
        # models -> geometry info tables + model info table
        models = np.zeros(max(1, len(scene.models)), dtype=abi.MODEL_INFO_DTYPE)
        self.gi_bufs = []
        for k, mo in enumerate(scene.models):
            pos, nrm, uv, idx = self.mesh_bufs[mo.mesh]
            mesh = scene.meshes[mo.mesh]
            gi = np.zeros(len(mesh.geometries), dtype=abi.GEOMETRY_INFO_DTYPE)
            for g in range(len(mesh.geometries)):
                d, m_, nm = mo.materials[g] if g < len(mo.materials) else mo.materials[-1]
                gi[g] = (idx[g].data_ptr(), d, m_, nm, 0)
            gbuf = _dev(gi, self.device)
            self.gi_bufs.append(gbuf)
            models[k] = (pos.data_ptr(), nrm.data_ptr(), uv.data_ptr(), gbuf.data_ptr())
        self.model_table = _dev(models, self.device)
        self.bindings = abi.SceneBindings()
        self.bindings.model_info = self.model_table.data_ptr()
        self.bindings.textures = self.tex_table.data_ptr()
        self.bindings.model_count = len(scene.models)
        self.bindings.texture_count = len(scene.textures)
        # instances + TLAS
        self.instances_host = scene.instance_array(self.blas)
        self.instances = _dev(self.instances_host, self.device)
        out = C.c_uint64()
        check(L.rtg_tlas_build(ctx.ptr, self.instances.data_ptr(), len(self.instances_host),
                               abi.RTG_BUILD_PREFER_FAST_TRACE | abi.RTG_BUILD_ALLOW_UPDATE, C.byref(out), sp),
              "rtg_tlas_build")
        self.tlas = out.value
        self.n_instances = len(self.instances_host)

    def instance_buffer(self, transforms):
        """A device instance buffer with new transforms (same BLAS refs)."""
        a = self.instances_host.copy()
        for k, xf in enumerate(transforms):
            a[k]["transform"] = xf
        return _dev(a, self.device)

    def refit(self, instance_buffer=None, tlas=None, stream=None):
        buf = self.instances if instance_buffer is None else instance_buffer
        check(load().rtg_tlas_refit(self.ctx.ptr, self.tlas if tlas is None else tlas, buf.data_ptr(),
                                    self.n_instances, _stream_ptr(stream)), "rtg_tlas_refit")

    def clone_tlas(self, stream=None):
        out = C.c_uint64()
        check(load().rtg_tlas_clone(self.ctx.ptr, self.tlas, C.byref(out), _stream_ptr(stream)), "rtg_tlas_clone")
        return out.value

    def as_bytes(self, address):
        """Raw bytes of an acceleration structure (synchronises)."""
        info = self.ctx.as_info(address)
        torch.cuda.synchronize()
        return device_to_host(address, info.size_bytes)

    def make_outputs(self, width=None, height=None, spp=None, hits=True):
        s = self.scene
        w = width or s.width
        h = height or s.height
        spp = spp or s.spp
        rgba = torch.zeros(w * h, dtype=torch.int32, device=self.device)
        rgb = torch.zeros(w * h * 3, dtype=torch.float32, device=self.device)
        ph = torch.zeros(w * h * spp * 6, dtype=torch.int32, device=self.device) if hits else None
        return rgba, rgb, ph

    def job(self, outputs, frame_index=None, opts=None, tlas=None, width=None, height=None, uniforms=None):
        s = self.scene
        rgba, rgb, ph = outputs
        j = abi.ImageJob()
        j.tlas = self.tlas if tlas is None else tlas
        j.scene = self.bindings
        j.uniforms = uniforms if uniforms is not None else s.uniforms(frame_index)
        j.width = width or s.width
        j.height = height or s.height
        j.opts = opts if opts is not None else s.trace_opts()
        j.out_rgba8 = rgba.data_ptr() if rgba is not None else 0
        j.out_rgb = rgb.data_ptr() if rgb is not None else 0
        j.out_primary_hits = ph.data_ptr() if ph is not None else 0
        return j

    def trace_image(self, frame_index=None, opts=None, outputs=None, stream=None, tlas=None):
        outputs = outputs or self.make_outputs(spp=(opts.spp if opts else None))
        j = self.job(outputs, frame_index, opts, tlas)
        check(load().rtg_trace_image(self.ctx.ptr, C.byref(j), _stream_ptr(stream)), "rtg_trace_image")
        return outputs

    def trace_rays(self, rays_dev, n, mode=abi.RTG_QUERY_CLOSEST, ray_flags=0, cull_mask=0xFF, sbt_offset=0,
                   out=None, stream=None, tlas=None):
        if out is None:
            if mode == abi.RTG_QUERY_CLOSEST:
                out = torch.zeros(n * 6, dtype=torch.int32, device=self.device)
            else:
                out = torch.zeros(n, dtype=torch.int32, device=self.device)
        hits = out.data_ptr() if mode == abi.RTG_QUERY_CLOSEST else 0
        occ = out.data_ptr() if mode == abi.RTG_QUERY_OCCLUSION else 0
        check(load().rtg_trace_rays(self.ctx.ptr, self.tlas if tlas is None else tlas, C.byref(self.bindings),
                                    rays_dev.data_ptr(), n, mode, ray_flags, cull_mask, sbt_offset, hits, occ,
                                    _stream_ptr(stream)), "rtg_trace_rays")
        return out

    def destroy(self):
        L = load()
        if getattr(self, "tlas", 0):
            L.rtg_tlas_destroy(self.ctx.ptr, self.tlas)
            self.tlas = 0
        for b in getattr(self, "blas", []):
            L.rtg_blas_destroy(self.ctx.ptr, b)
        self.blas = []


_hip = None


def device_to_host(address, nbytes):
    """hipMemcpy of a raw device allocation into a new numpy uint8 array."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    out = np.empty(nbytes, dtype=np.uint8)
    st = _hip.hipMemcpy(out.ctypes.data, C.c_void_p(address), nbytes, 2)  # hipMemcpyDeviceToHost
    if st != 0:
        raise RuntimeError(f"hipMemcpy failed: {st}")
    return out


def hits_to_numpy(t):
    """(N*6) int32 device tensor -> structured HIT_DTYPE array."""
    return t.cpu().numpy().view(abi.HIT_DTYPE)


def rays_to_device(rays, device="cuda:0"):
    return _dev(np.ascontiguousarray(rays, dtype=abi.RAY_DTYPE), torch.device(device))
