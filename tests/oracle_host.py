"""Host-side driver of the CPU oracle (oracle/_build/liborc.so).

Test infrastructure: used only by tests, __graft_entry__.smoke() and the CPU
baseline leg of bench.py, as the checker. Device addresses of the ABI structs
are host pointers of numpy arrays kept alive by HostScene.
"""
import ctypes as C
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "ray-tracing-gallery_amd"))

from rtg import abi  # noqa: E402

ORC_PATH = os.path.join(REPO, "oracle", "_build", "liborc.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORC_PATH):
            raise RuntimeError(f"oracle not built: {ORC_PATH} (run make -C oracle)")
        L = C.CDLL(ORC_PATH)
        vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
        L.orc_blas_build.argtypes = [vp, u32, u32, C.POINTER(vp), C.POINTER(u64)]
        L.orc_tlas_build.argtypes = [vp, u32, u32, C.POINTER(vp), C.POINTER(u64)]
        L.orc_debug_binary_tree.argtypes = [vp, u32, u32, vp, vp, vp]
        L.orc_tlas_refit.argtypes = [vp, vp, u32]
        L.orc_free.argtypes = [vp]
        L.orc_as_get_info.argtypes = [vp, C.POINTER(abi.AsInfo)]
        L.orc_trace_rays.argtypes = [vp, vp, vp, u32, u32, u32, u32, u32, vp, vp, C.POINTER(abi.TraceCounters), i32]
        L.orc_trace_rays_brute.argtypes = [vp, vp, vp, u32, u32, u32, u32, u32, vp, vp]
        L.orc_set_stack_lds.argtypes = [u32]
        L.orc_set_stack_lds.restype = None
        L.orc_set_shadow_pairs.argtypes = [C.c_int]
        L.orc_set_shadow_pairs.restype = None
        L.orc_debug_visits.argtypes = [vp, vp, vp, vp, vp, i32, vp, u32, vp]
        L.orc_trace_image.argtypes = [C.POINTER(abi.ImageJob), vp, u32, i32, C.POINTER(abi.TraceCounters)]
        L.orc_v_smith_ggx_correlated.argtypes = [C.c_float] * 3
        L.orc_v_smith_ggx_correlated.restype = C.c_float
        L.orc_d_ggx.argtypes = [C.c_float] * 2
        L.orc_d_ggx.restype = C.c_float
        fp = C.POINTER(C.c_float)
        L.orc_brdf.argtypes = [fp, fp, fp, C.c_float, fp, C.c_float, C.c_float, C.c_float, fp]
        L.orc_linear_to_srgb.argtypes = [fp, fp]
        L.orc_heatmap_temperature.argtypes = [C.c_float, fp]
        L.orc_sample_directional_light.argtypes = [fp, fp, C.c_float, fp]
        L.orc_sample_texture.argtypes = [C.POINTER(abi.Texture), fp, C.c_float, C.c_float, fp]
        L.orc_srgb_lut.argtypes = [fp]
        for name in ("orc_cos", "orc_sin", "orc_log2", "orc_exp2"):
            getattr(L, name).argtypes = [C.c_float]
            getattr(L, name).restype = C.c_float
        L.orc_pow.argtypes = [C.c_float, C.c_float]
        L.orc_pow.restype = C.c_float
        L.orc_watertight.argtypes = [fp, fp, fp, fp, fp, C.c_float, C.c_float, C.POINTER(C.c_int), fp, fp, fp]
        L.orc_morton3.argtypes = [C.c_float] * 3 + [fp, fp]
        L.orc_morton3.restype = C.c_uint32
        L.orc_camera_ray.argtypes = [C.POINTER(abi.Uniforms), u32, u32, u32, u32, C.c_float, C.c_float, fp, fp]
        L.orc_primary_miss.argtypes = [fp, fp, C.c_float, fp]
        L.orc_terminator_fix.argtypes = [fp, fp, fp, fp]
        L.orc_blue_noise.argtypes = [C.POINTER(abi.Texture), fp, u32, u32, u32, u32, fp]
        L.orc_shading_normal.argtypes = [fp, fp, fp, fp, fp, fp]
        L.orc_textured_colour.argtypes = [fp, fp, fp, C.c_float, fp, C.c_float, C.c_float, fp]
        L.orc_mirror.argtypes = [fp, fp, fp, fp, C.c_float, fp, fp]
        L.orc_portal.argtypes = [fp, fp, C.c_float, fp, fp]
        L.orc_heatmap_pixels.argtypes = [vp, vp, u32, vp]
        for name in ("orc_camera_ray", "orc_primary_miss", "orc_terminator_fix", "orc_blue_noise",
                     "orc_shading_normal", "orc_textured_colour", "orc_mirror", "orc_portal", "orc_heatmap_pixels"):
            getattr(L, name).restype = None
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class HostBlob:
    def __init__(self, ptr, size):
        self.ptr = ptr
        self.size = size

    def bytes(self):
        return C.string_at(self.ptr, self.size)

    def info(self):
        inf = abi.AsInfo()
        assert lib().orc_as_get_info(self.ptr, C.byref(inf)) == 0
        return inf

    def __del__(self):
        if self.ptr and _lib is not None:
            _lib.orc_free(self.ptr)
            self.ptr = None


def geometries_for(mesh, keep):
    """rtg_geometry descriptors (host pointers) of a mesh's geometries."""
    arr = (abi.Geometry * max(1, len(mesh.geometries)))()
    pos = np.ascontiguousarray(mesh.positions, dtype=np.float32)
    keep.append(pos)
    for g, (ix, op) in enumerate(mesh.geometries):
        ix = np.ascontiguousarray(ix, dtype=np.uint32)
        keep.append(ix)
        arr[g].vertex_data = _ptr(pos)
        arr[g].vertex_stride = 12
        arr[g].max_vertex = max(0, len(pos) - 1)
        arr[g].index_data = _ptr(ix)
        arr[g].index_count = ix.size
        arr[g].flags = abi.RTG_GEOMETRY_OPAQUE if op else 0
    return arr


# the flags rtg/device.py builds with (the reference's PREFER_FAST_TRACE)
BLAS_FLAGS = abi.RTG_BUILD_PREFER_FAST_TRACE
TLAS_FLAGS = abi.RTG_BUILD_PREFER_FAST_TRACE | abi.RTG_BUILD_ALLOW_UPDATE


def build_blas(mesh, keep=None, flags=BLAS_FLAGS):
    keep = [] if keep is None else keep
    arr = geometries_for(mesh, keep)
    out = C.c_void_p()
    size = C.c_uint64()
    st = lib().orc_blas_build(C.cast(arr, C.c_void_p), len(mesh.geometries), flags, C.byref(out), C.byref(size))
    if st != 0:
        raise RuntimeError(f"orc_blas_build failed: {st}")
    return HostBlob(out.value, size.value)


class HostScene:
    """A Scene's buffers, descriptor tables and acceleration structures in host
    memory for the oracle."""

    def __init__(self, scene, blas_flags=BLAS_FLAGS, tlas_flags=TLAS_FLAGS):
        L = lib()
        self.scene = scene
        self.keep = []
        self.blas = [build_blas(m, self.keep, blas_flags) for m in scene.meshes]
        # textures
        self.tex_table = np.zeros(max(1, len(scene.textures)), dtype=abi.TEXTURE_DTYPE)
        for k, t in enumerate(scene.textures):
            data = np.ascontiguousarray(t.data)
            self.keep.append(data)
            self.tex_table[k] = (_ptr(data), t.width, t.height, t.format, t.filter)
        # models
        self.model_table = np.zeros(max(1, len(scene.models)), dtype=abi.MODEL_INFO_DTYPE)
        for k, mo in enumerate(scene.models):
            mesh = scene.meshes[mo.mesh]
            gi = np.zeros(len(mesh.geometries), dtype=abi.GEOMETRY_INFO_DTYPE)
            for g, (ix, _) in enumerate(mesh.geometries):
                ix = np.ascontiguousarray(ix, dtype=np.uint32)
                self.keep.append(ix)
                d, m, nm = mo.materials[g] if g < len(mo.materials) else mo.materials[-1]
                gi[g] = (_ptr(ix), d, m, nm, 0)
            self.keep.append(gi)
            for arr in (mesh.positions, mesh.normals, mesh.uvs):
                self.keep.append(arr)
            self.model_table[k] = (_ptr(mesh.positions), _ptr(mesh.normals), _ptr(mesh.uvs), _ptr(gi))
        self.bindings = abi.SceneBindings()
        self.bindings.model_info = _ptr(self.model_table)
        self.bindings.textures = _ptr(self.tex_table)
        self.bindings.model_count = len(scene.models)
        self.bindings.texture_count = len(scene.textures)
        self.instances = scene.instance_array([b.ptr for b in self.blas])
        out = C.c_void_p()
        size = C.c_uint64()
        st = L.orc_tlas_build(_ptr(self.instances), len(self.instances), tlas_flags, C.byref(out), C.byref(size))
        if st != 0:
            raise RuntimeError(f"orc_tlas_build failed: {st}")
        self.tlas = HostBlob(out.value, size.value)

    def refit(self, transforms=None):
        """Writes new transforms into the instance records and refits."""
        if transforms is not None:
            for k, xf in enumerate(transforms):
                self.instances[k]["transform"] = xf
        st = lib().orc_tlas_refit(self.tlas.ptr, _ptr(self.instances), len(self.instances))
        if st != 0:
            raise RuntimeError(f"orc_tlas_refit failed: {st}")

    def job(self, frame_index=None, want_hits=True, opts=None):
        s = self.scene
        n = s.width * s.height
        spp = (opts.spp if opts else s.spp)
        rgba = np.zeros(n, dtype=np.uint32)
        rgb = np.zeros(n * 3, dtype=np.float32)
        hits = np.zeros(n * spp, dtype=abi.HIT_DTYPE) if want_hits else None
        j = abi.ImageJob()
        j.tlas = self.tlas.ptr
        j.scene = self.bindings
        j.uniforms = s.uniforms(frame_index)
        j.width, j.height = s.width, s.height
        j.opts = opts if opts is not None else s.trace_opts()
        j.out_rgba8 = _ptr(rgba)
        j.out_rgb = _ptr(rgb)
        j.out_primary_hits = _ptr(hits) if want_hits else 0
        return j, rgba, rgb, hits

    def trace_image(self, frame_index=None, pixels=None, threads=0, want_hits=True, opts=None):
        j, rgba, rgb, hits = self.job(frame_index, want_hits, opts)
        cnt = abi.TraceCounters()
        if pixels is not None:
            pixels = np.ascontiguousarray(pixels, dtype=np.uint32)
            st = lib().orc_trace_image(C.byref(j), _ptr(pixels), len(pixels), threads, C.byref(cnt))
        else:
            st = lib().orc_trace_image(C.byref(j), None, 0, threads, C.byref(cnt))
        if st != 0:
            raise RuntimeError(f"orc_trace_image failed: {st}")
        return rgba, rgb.reshape(-1, 3), hits, cnt

    def trace_rays(self, rays, mode=abi.RTG_QUERY_CLOSEST, ray_flags=0, cull_mask=0xFF, sbt_offset=0, threads=0,
                   brute=False):
        rays = np.ascontiguousarray(rays, dtype=abi.RAY_DTYPE)
        n = len(rays)
        hits = np.zeros(n, dtype=abi.HIT_DTYPE)
        occ = np.zeros(n, dtype=np.uint32)
        if brute:
            st = lib().orc_trace_rays_brute(self.tlas.ptr, C.byref(self.bindings), _ptr(rays), n, mode, ray_flags,
                                            cull_mask, sbt_offset, _ptr(hits), _ptr(occ))
            cnt = None
        else:
            cnt = abi.TraceCounters()
            st = lib().orc_trace_rays(self.tlas.ptr, C.byref(self.bindings), _ptr(rays), n, mode, ray_flags,
                                      cull_mask, sbt_offset, _ptr(hits), _ptr(occ), C.byref(cnt), threads)
        if st != 0:
            raise RuntimeError(f"orc_trace_rays failed: {st}")
        return (hits if mode == abi.RTG_QUERY_CLOSEST else occ), cnt


def binary_tree(boxes, flags=BLAS_FLAGS):
    """The binary tree (lc, rc, order) the oracle's build over `boxes` ((n, 6)
    float32: lo xyz, hi xyz; n >= 2) collapses into wide nodes."""
    b = np.ascontiguousarray(boxes, dtype=np.float32)
    n = len(b)
    lc = np.zeros(n - 1, np.uint32)
    rc = np.zeros(n - 1, np.uint32)
    order = np.zeros(n, np.uint32)
    st = lib().orc_debug_binary_tree(_ptr(b), n, flags, _ptr(lc), _ptr(rc), _ptr(order))
    if st != 0:
        raise RuntimeError(f"orc_debug_binary_tree failed: {st}")
    return lc, rc, order
