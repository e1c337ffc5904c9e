"""The host frame driver (rtg_frames_*): frames in flight over one scene, as
the reference's frame loop runs them (src/main.rs:600-668, 916-996): wait for
the frame's slot, write the instance transforms into its mapped array, then
upload, refit the slot's TLAS copy and trace."""
import ctypes as C

import numpy as np

from . import abi
from .device import _stream_ptr
from .lib import check, load


class FrameDriver:
    def __init__(self, scene_dev, frames_in_flight=2):
        self.ds = scene_dev
        self.n = len(scene_dev.scene.instances)
        self.ptr = C.c_void_p()
        check(load().rtg_frames_create(scene_dev.ctx.ptr, scene_dev.tlas, frames_in_flight, C.byref(self.ptr)),
              "rtg_frames_create")

    def begin(self, frame):
        """Waits for the frame's slot; returns (instances, slot TLAS), where
        instances is a writable numpy view (abi.INSTANCE_DTYPE) of the slot's
        pinned host array."""
        p = C.POINTER(abi.Instance)()
        tl = C.c_uint64()
        check(load().rtg_frames_begin(self.ptr, frame, C.byref(p), C.byref(tl)), "rtg_frames_begin")
        raw = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(max(1, self.n) * 64,))
        return raw[:self.n * 64].view(abi.INSTANCE_DTYPE), int(tl.value)

    def submit(self, frame, job, stream=None):
        check(load().rtg_frames_submit(self.ptr, frame, C.byref(job), _stream_ptr(stream)), "rtg_frames_submit")

    def close(self):
        if self.ptr:
            check(load().rtg_frames_destroy(self.ptr), "rtg_frames_destroy")
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
