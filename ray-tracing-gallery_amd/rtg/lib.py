"""ctypes binding of librtg.so, the C ABI of include/rtg.h.

The library is the product: there is no fallback. Loading fails loudly when
the HIP build is missing.
"""
import ctypes as C
import os

from . import abi

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "librtg.so")

# Every extern "C" symbol of include/rtg.h.
EXPORTS = [
    "rtg_create", "rtg_destroy", "rtg_last_error", "rtg_abi_version",
    "rtg_blas_query_sizes", "rtg_blas_build", "rtg_blas_destroy",
    "rtg_tlas_query_sizes", "rtg_tlas_build", "rtg_tlas_refit", "rtg_tlas_clone", "rtg_tlas_destroy",
    "rtg_as_get_info", "rtg_trace_rays", "rtg_trace_rays_static", "rtg_trace_image", "rtg_trace_batch",
    "rtg_set_counters", "rtg_get_counters", "rtg_get_status", "rtg_set_timing", "rtg_get_kernel_time",
    "rtg_server_create", "rtg_server_add_scene", "rtg_server_submit", "rtg_server_cancel", "rtg_server_poll",
    "rtg_server_replay",
    "rtg_server_clock", "rtg_server_get_stats", "rtg_server_destroy",
    "rtg_get_tuning", "rtg_set_tuning", "rtg_autotune",
    "rtg_batch_create", "rtg_batch_update", "rtg_batch_launch", "rtg_batch_destroy",
    "rtg_frames_create", "rtg_frames_begin", "rtg_frames_submit", "rtg_frames_destroy",
]

_lib = None


class RtgError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RtgError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build() or make -C "
                       f"ray-tracing-gallery_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    L.rtg_create.argtypes = [C.POINTER(vp), i32]
    L.rtg_destroy.argtypes = [vp]
    L.rtg_destroy.restype = None
    L.rtg_last_error.restype = C.c_char_p
    L.rtg_abi_version.restype = u32
    L.rtg_blas_query_sizes.argtypes = [vp, vp, u32, C.POINTER(abi.BuildSizes)]
    L.rtg_blas_build.argtypes = [vp, vp, u32, u32, C.POINTER(u64), vp]
    L.rtg_blas_destroy.argtypes = [vp, u64]
    L.rtg_tlas_query_sizes.argtypes = [vp, u32, C.POINTER(abi.BuildSizes)]
    L.rtg_tlas_build.argtypes = [vp, u64, u32, u32, C.POINTER(u64), vp]
    L.rtg_tlas_refit.argtypes = [vp, u64, u64, u32, vp]
    L.rtg_tlas_clone.argtypes = [vp, u64, C.POINTER(u64), vp]
    L.rtg_tlas_destroy.argtypes = [vp, u64]
    L.rtg_as_get_info.argtypes = [vp, u64, C.POINTER(abi.AsInfo)]
    L.rtg_trace_rays.argtypes = [vp, u64, C.POINTER(abi.SceneBindings), u64, u32, u32, u32, u32, u32, u64, u64, vp]
    L.rtg_trace_rays_static.argtypes = L.rtg_trace_rays.argtypes
    L.rtg_trace_image.argtypes = [vp, C.POINTER(abi.ImageJob), vp]
    L.rtg_trace_batch.argtypes = [vp, vp, u32, vp]
    L.rtg_set_counters.argtypes = [vp, i32]
    L.rtg_get_counters.argtypes = [vp, C.POINTER(abi.TraceCounters)]
    L.rtg_get_status.argtypes = [vp, C.POINTER(u32), i32]
    L.rtg_set_timing.argtypes = [vp, i32]
    L.rtg_get_kernel_time.argtypes = [vp, u32, C.POINTER(C.c_double), C.POINTER(u64)]
    L.rtg_server_create.argtypes = [vp, C.POINTER(abi.ServerConfig), C.POINTER(vp)]
    L.rtg_server_add_scene.argtypes = [vp, u64, C.POINTER(abi.SceneBindings), C.POINTER(u32)]
    L.rtg_server_submit.argtypes = [vp, C.POINTER(abi.Request), C.POINTER(u64)]
    L.rtg_server_cancel.argtypes = [vp, u64]
    L.rtg_server_poll.argtypes = [vp, C.POINTER(abi.Completion), u32, C.POINTER(u32), i32]
    L.rtg_server_replay.argtypes = [vp, C.POINTER(abi.Request), C.POINTER(C.c_double), u32,
                                    C.POINTER(abi.Completion)]
    L.rtg_server_clock.argtypes = [vp]
    L.rtg_server_get_stats.argtypes = [vp, C.POINTER(abi.ServerStats)]
    L.rtg_server_destroy.argtypes = [vp]
    L.rtg_get_tuning.argtypes = [vp, C.POINTER(abi.Tuning)]
    L.rtg_set_tuning.argtypes = [vp, C.POINTER(abi.Tuning)]
    L.rtg_autotune.argtypes = [vp, vp, u32, u32, C.POINTER(abi.Tuning), C.POINTER(C.c_double)]
    L.rtg_batch_create.argtypes = [vp, vp, u32, C.POINTER(vp)]
    L.rtg_batch_update.argtypes = [vp, vp, u32, vp]
    L.rtg_batch_launch.argtypes = [vp, vp]
    L.rtg_batch_destroy.argtypes = [vp]
    L.rtg_frames_create.argtypes = [vp, u64, u32, C.POINTER(vp)]
    L.rtg_frames_begin.argtypes = [vp, u64, C.POINTER(C.POINTER(abi.Instance)), C.POINTER(u64)]
    L.rtg_frames_submit.argtypes = [vp, u64, C.POINTER(abi.ImageJob), vp]
    L.rtg_frames_destroy.argtypes = [vp]
    for name in EXPORTS:
        if name not in ("rtg_destroy", "rtg_last_error", "rtg_abi_version"):
            getattr(L, name).restype = C.c_int
    L.rtg_server_clock.restype = C.c_double
    _lib = L
    return L


def check(status, what=""):
    if status != 0:
        msg = load().rtg_last_error().decode(errors="replace")
        raise RtgError(f"{what} failed ({status}): {msg}")


class Context:
    """Owns an rtg_ctx (one per device)."""

    def __init__(self, device=0):
        L = load()
        self.ptr = C.c_void_p()
        check(L.rtg_create(C.byref(self.ptr), device), "rtg_create")
        self.device = device

    def close(self):
        if self.ptr:
            load().rtg_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_counters(self, on):
        check(load().rtg_set_counters(self.ptr, 1 if on else 0), "rtg_set_counters")

    def counters(self):
        c = abi.TraceCounters()
        check(load().rtg_get_counters(self.ptr, C.byref(c)), "rtg_get_counters")
        return c

    def status(self, clear=True):
        """Sticky device status bits (abi.RTG_STATUS_*); synchronises."""
        v = C.c_uint32()
        check(load().rtg_get_status(self.ptr, C.byref(v), 1 if clear else 0), "rtg_get_status")
        return int(v.value)

    def set_timing(self, on):
        check(load().rtg_set_timing(self.ptr, 1 if on else 0), "rtg_set_timing")

    def kernel_time(self, kind):
        ms = C.c_double()
        n = C.c_uint64()
        check(load().rtg_get_kernel_time(self.ptr, kind, C.byref(ms), C.byref(n)), "rtg_get_kernel_time")
        return ms.value, n.value

    def tuning(self):
        t = abi.Tuning()
        check(load().rtg_get_tuning(self.ptr, C.byref(t)), "rtg_get_tuning")
        return t

    def set_tuning(self, t):
        check(load().rtg_set_tuning(self.ptr, C.byref(t)), "rtg_set_tuning")

    def as_info(self, address):
        inf = abi.AsInfo()
        check(load().rtg_as_get_info(self.ptr, address, C.byref(inf)), "rtg_as_get_info")
        return inf
