"""Batched launches: graph-captured batches (rtg_batch_*), the eager batch
(rtg_trace_batch) and the autotuner (rtg_autotune) of include/rtg.h."""
import ctypes as C

from . import abi
from .lib import check, load
from .device import _stream_ptr


def job_array(jobs):
    return (abi.ImageJob * len(jobs))(*jobs)


def trace_batch(ctx, jobs, stream=None):
    """All jobs in one set of persistent launches (no graph)."""
    arr = job_array(jobs)
    check(load().rtg_trace_batch(ctx.ptr, C.cast(arr, C.c_void_p), len(jobs), _stream_ptr(stream)),
          "rtg_trace_batch")


def autotune(ctx, jobs, repeats=2):
    """Searches the launch settings on `jobs`; the best becomes the context's
    setting. Returns (abi.Tuning, median ms)."""
    arr = job_array(jobs)
    best = abi.Tuning()
    ms = C.c_double()
    check(load().rtg_autotune(ctx.ptr, C.cast(arr, C.c_void_p), len(jobs), repeats, C.byref(best), C.byref(ms)),
          "rtg_autotune")
    return best, ms.value


class Batch:
    """A fixed set of image jobs captured into one HIP graph."""

    def __init__(self, ctx, jobs):
        self.ctx = ctx
        self.n = len(jobs)
        self._jobs = job_array(jobs)
        self.ptr = C.c_void_p()
        check(load().rtg_batch_create(ctx.ptr, C.cast(self._jobs, C.c_void_p), self.n, C.byref(self.ptr)),
              "rtg_batch_create")

    def update(self, jobs, stream=None):
        self._jobs = job_array(jobs)
        check(load().rtg_batch_update(self.ptr, C.cast(self._jobs, C.c_void_p), len(jobs), _stream_ptr(stream)),
              "rtg_batch_update")

    def launch(self, stream=None):
        check(load().rtg_batch_launch(self.ptr, _stream_ptr(stream)), "rtg_batch_launch")

    def close(self):
        if self.ptr:
            check(load().rtg_batch_destroy(self.ptr), "rtg_batch_destroy")
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
