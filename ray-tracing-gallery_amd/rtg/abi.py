"""ctypes mirrors of the C ABI structs declared in include/rtg.h.

The layouts are the reference's host/device ABI:
  rtg_instance      vk::AccelerationStructureInstanceKHR twin (src/gpu_structs.rs:17-58)
  rtg_uniforms      Uniforms (shared-structs/src/lib.rs:7-19)
  rtg_model_info    ModelInfo (shared-structs/src/lib.rs:21-29)
  rtg_geometry_info GeometryInfo + GeometryImages (shared-structs/src/lib.rs:31-47)
"""
import ctypes as C

import numpy as np

RTG_OK = 0
RTG_NO_HIT = 0xFFFFFFFF

RTG_INSTANCE_TRIANGLE_FACING_CULL_DISABLE = 0x1
RTG_INSTANCE_FORCE_OPAQUE = 0x4
RTG_INSTANCE_FORCE_NO_OPAQUE = 0x8

RTG_HIT_TEXTURED = 0
RTG_HIT_MIRROR = 1
RTG_HIT_PORTAL = 2
RTG_HIT_GROUP_COUNT = 3

RTG_FORMAT_RGBA8_UNORM = 0
RTG_FORMAT_RGBA8_SRGB = 1
RTG_FORMAT_RGBA32_FLOAT = 2
RTG_FILTER_NEAREST = 0
RTG_FILTER_LINEAR = 1

RTG_GEOMETRY_OPAQUE = 0x1
RTG_BUILD_ALLOW_UPDATE = 0x1
RTG_BUILD_PREFER_FAST_TRACE = 0x4

RTG_RAY_FLAG_NONE = 0x0
RTG_RAY_FLAG_OPAQUE = 0x1
RTG_RAY_FLAG_NO_OPAQUE = 0x2
RTG_RAY_FLAG_TERMINATE_ON_FIRST_HIT = 0x4
RTG_RAY_FLAG_SKIP_CLOSEST_HIT_SHADER = 0x8

RTG_QUERY_CLOSEST = 0
RTG_QUERY_OCCLUSION = 1

RTG_TRACE_HEATMAP = 0x1
RTG_TRACE_OUTPUT_BGRA8 = 0x2

RTG_STATUS_BAD_INDEX = 0x1
RTG_STATUS_UNKNOWN_BLAS = 0x2
RTG_STATUS_STACK_OVERFLOW = 0x4
RTG_STATUS_QUEUE_RANGE = 0x8

RTG_KERNEL_NAMES = ["raygen", "trace_closest", "shade", "trace_occlusion", "resolve", "refit", "batch_graph",
                    "bounce_side"]
RTG_KERNEL_RAYGEN = 0
RTG_KERNEL_TRACE_CLOSEST = 1
RTG_KERNEL_SHADE = 2
RTG_KERNEL_TRACE_OCCLUSION = 3
RTG_KERNEL_RESOLVE = 4
RTG_KERNEL_REFIT = 5
RTG_KERNEL_BATCH_GRAPH = 6
RTG_KERNEL_BOUNCE_SIDE = 7


class Instance(C.Structure):
    _fields_ = [
        ("transform", (C.c_float * 4) * 3),
        ("custom_index_and_mask", C.c_uint32),
        ("sbt_offset_and_flags", C.c_uint32),
        ("acceleration_structure_reference", C.c_uint64),
    ]


class Uniforms(C.Structure):
    _fields_ = [
        ("view_inverse", C.c_float * 16),
        ("proj_inverse", C.c_float * 16),
        ("sun_dir", C.c_float * 3),
        ("_padding0", C.c_uint32),
        ("sun_radius", C.c_float),
        ("blue_noise_texture_index", C.c_uint32),
        ("ggx_lut_texture_index", C.c_uint32),
        ("frame_index", C.c_uint32),
        ("show_heatmap", C.c_uint8),
        ("_padding1", C.c_uint8 * 15),
    ]


class ModelInfo(C.Structure):
    _fields_ = [
        ("position_buffer_address", C.c_uint64),
        ("normal_buffer_address", C.c_uint64),
        ("uv_buffer_address", C.c_uint64),
        ("geometry_info_address", C.c_uint64),
    ]


class GeometryInfo(C.Structure):
    _fields_ = [
        ("index_buffer_address", C.c_uint64),
        ("diffuse_image_index", C.c_uint32),
        ("metallic_roughness_image_index", C.c_uint32),
        ("normal_map_image_index", C.c_int32),
        ("_padding", C.c_uint32),
    ]


class Texture(C.Structure):
    _fields_ = [
        ("data", C.c_uint64),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("format", C.c_uint32),
        ("filter", C.c_uint32),
    ]


class Geometry(C.Structure):
    _fields_ = [
        ("vertex_data", C.c_uint64),
        ("vertex_stride", C.c_uint32),
        ("max_vertex", C.c_uint32),
        ("index_data", C.c_uint64),
        ("index_count", C.c_uint32),
        ("flags", C.c_uint32),
    ]


class BuildSizes(C.Structure):
    _fields_ = [
        ("acceleration_structure_size", C.c_uint64),
        ("build_scratch_size", C.c_uint64),
        ("update_scratch_size", C.c_uint64),
    ]


class AsInfo(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("node_count", C.c_uint32),
        ("prim_count", C.c_uint32),
        ("geometry_count", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("node_width", C.c_uint32),
        ("bounds", C.c_float * 6),
        ("size_bytes", C.c_uint64),
    ]


class Ray(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 3),
        ("tmin", C.c_float),
        ("direction", C.c_float * 3),
        ("tmax", C.c_float),
    ]


class Hit(C.Structure):
    _fields_ = [
        ("t", C.c_float),
        ("instance", C.c_uint32),
        ("primitive", C.c_uint32),
        ("geometry", C.c_uint32),
        ("u", C.c_float),
        ("v", C.c_float),
    ]


class SceneBindings(C.Structure):
    _fields_ = [
        ("model_info", C.c_uint64),
        ("textures", C.c_uint64),
        ("model_count", C.c_uint32),
        ("texture_count", C.c_uint32),
    ]


class TraceOpts(C.Structure):
    _fields_ = [
        ("shadow_rays", C.c_uint32),
        ("max_traces", C.c_uint32),
        ("spp", C.c_uint32),
        ("flags", C.c_uint32),
    ]


class ImageJob(C.Structure):
    _fields_ = [
        ("tlas", C.c_uint64),
        ("scene", SceneBindings),
        ("uniforms", Uniforms),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("opts", TraceOpts),
        ("out_rgba8", C.c_uint64),
        ("out_rgb", C.c_uint64),
        ("out_primary_hits", C.c_uint64),
    ]


class TraceCounters(C.Structure):
    _fields_ = [
        ("primary_rays", C.c_uint64),
        ("bounce_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("node_visits", C.c_uint64),
        ("triangle_tests", C.c_uint64),
        ("instance_visits", C.c_uint64),
        ("anyhit_calls", C.c_uint64),
        ("occ_node_visits", C.c_uint64),
        ("occ_triangle_tests", C.c_uint64),
        ("occ_instance_visits", C.c_uint64),
        ("occ_anyhit_calls", C.c_uint64),
        ("side_shadow_rays", C.c_uint64),
        ("side_occ_node_visits", C.c_uint64),
        ("side_occ_triangle_tests", C.c_uint64),
        ("side_occ_instance_visits", C.c_uint64),
        ("side_occ_anyhit_calls", C.c_uint64),
        ("side_node_visits", C.c_uint64),
        ("side_triangle_tests", C.c_uint64),
        ("side_instance_visits", C.c_uint64),
        ("side_anyhit_calls", C.c_uint64),
        ("spill_pushes", C.c_uint64),
        ("occ_spill_pushes", C.c_uint64),
        ("side_spill_pushes", C.c_uint64),
        ("side_occ_spill_pushes", C.c_uint64),
        ("textured_hits", C.c_uint64),
        ("side_textured_hits", C.c_uint64),
        ("side_bounce_rays", C.c_uint64),
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


# numpy views of the record arrays (same byte layout as the C structs)
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("tmin", "<f4"), ("direction", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("instance", "<u4"), ("primitive", "<u4"), ("geometry", "<u4"),
                      ("u", "<f4"), ("v", "<f4")])
INSTANCE_DTYPE = np.dtype([("transform", "<f4", (3, 4)), ("custom_index_and_mask", "<u4"),
                           ("sbt_offset_and_flags", "<u4"), ("blas", "<u8")])
MODEL_INFO_DTYPE = np.dtype([("position", "<u8"), ("normal", "<u8"), ("uv", "<u8"), ("geometry_info", "<u8")])
GEOMETRY_INFO_DTYPE = np.dtype([("index", "<u8"), ("diffuse", "<u4"), ("metallic_roughness", "<u4"),
                                ("normal_map", "<i4"), ("_padding", "<u4")])
TEXTURE_DTYPE = np.dtype([("data", "<u8"), ("width", "<u4"), ("height", "<u4"), ("format", "<u4"),
                          ("filter", "<u4")])

assert C.sizeof(Instance) == 64 and INSTANCE_DTYPE.itemsize == 64
assert C.sizeof(Uniforms) == 176
assert C.sizeof(ModelInfo) == 32 and MODEL_INFO_DTYPE.itemsize == 32
assert C.sizeof(GeometryInfo) == 24 and GEOMETRY_INFO_DTYPE.itemsize == 24
assert C.sizeof(Texture) == 24 and TEXTURE_DTYPE.itemsize == 24
assert C.sizeof(Ray) == 32 and RAY_DTYPE.itemsize == 32
assert C.sizeof(Hit) == 24 and HIT_DTYPE.itemsize == 24


def pack24_8(low24, high8):
    """vk::Packed24_8::new (src/gpu_structs.rs:45-49)."""
    return (int(low24) & 0xFFFFFF) | ((int(high8) & 0xFF) << 24)


def uniforms_from(view_inverse, proj_inverse, sun_dir, sun_radius, blue_noise_index, lut_index, frame_index,
                  show_heatmap=False):
    """Fills a Uniforms struct; matrices are 4x4 numpy arrays in math (row, col)
    order and are stored column-major as glam's Mat4."""
    u = Uniforms()
    vi = np.asarray(view_inverse, dtype=np.float32).T.reshape(-1)
    pi = np.asarray(proj_inverse, dtype=np.float32).T.reshape(-1)
    for i in range(16):
        u.view_inverse[i] = float(vi[i])
        u.proj_inverse[i] = float(pi[i])
    for i in range(3):
        u.sun_dir[i] = float(np.float32(sun_dir[i]))
    u.sun_radius = float(np.float32(sun_radius))
    u.blue_noise_texture_index = blue_noise_index
    u.ggx_lut_texture_index = lut_index
    u.frame_index = frame_index
    u.show_heatmap = 1 if show_heatmap else 0
    return u


class Tuning(C.Structure):
    _fields_ = [
        ("blocks_per_cu", C.c_uint32),
        ("feed_chunk", C.c_uint32),
        ("refill_closest", C.c_uint32),
        ("refill_occlusion", C.c_uint32),
        ("path_order", C.c_uint32),
        ("bounce_blocks", C.c_uint32),
        ("stack_lds_entries", C.c_uint32),
        ("reserved", C.c_uint32),
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_ if name != "reserved"}

    @staticmethod
    def make(blocks_per_cu=4, feed_chunk=64, refill_closest=64, refill_occlusion=64, path_order=1,
             bounce_blocks=192, stack_lds_entries=16):
        return Tuning(blocks_per_cu, feed_chunk, refill_closest, refill_occlusion, path_order, bounce_blocks,
                      stack_lds_entries)


assert C.sizeof(Tuning) == 32


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch_jobs", C.c_uint32), ("max_inflight", C.c_uint32),
                ("max_batch_samples", C.c_uint64)]


assert C.sizeof(ServerConfig) == 24


class Request(C.Structure):
    _fields_ = [
        ("scene", C.c_uint32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("opts", TraceOpts),
        ("uniforms", Uniforms),
        ("instances", C.c_uint64),
        ("out_rgba8", C.c_uint64),
        ("out_rgb", C.c_uint64),
        ("out_primary_hits", C.c_uint64),
        ("tag", C.c_uint64),
        ("deadline_s", C.c_double),
    ]


class Completion(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64),
        ("tag", C.c_uint64),
        ("status", C.c_int32),
        ("batch_size", C.c_uint32),
        ("submit_s", C.c_double),
        ("start_s", C.c_double),
        ("done_s", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("batches", C.c_uint64), ("rejected", C.c_uint64),
                ("max_batch", C.c_uint64), ("max_inflight_seen", C.c_uint64), ("peak_bytes", C.c_uint64)]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


RTG_OK = 0
RTG_ERROR_INVALID_ARGUMENT = -1
RTG_ERROR_OUT_OF_MEMORY = -2
RTG_ERROR_DEVICE = -3
RTG_ERROR_LIMIT = -4
RTG_ERROR_NOT_FOUND = -5
RTG_ERROR_BUDGET = -6
RTG_ERROR_STATE = -7
RTG_ERROR_RESULT = -8
RTG_ERROR_CANCELLED = -9
RTG_ERROR_TIMEOUT = -10
