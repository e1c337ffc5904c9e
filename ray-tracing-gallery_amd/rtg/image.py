"""The output step: read a traced image back and write it as a PNG.

The reference shows its storage image by copying it to the swapchain
(src/command_buffer_recording.rs:128-207); a service returns it to a caller
instead. `readback` brings a packed RGBA8 or BGRA8 image (one u32 per pixel,
as rtg_image_job.out_rgba8 holds it) to host memory as an (H, W, 4) RGBA byte
array; `write_png` encodes that array as an 8-bit RGBA PNG (zlib, no filter),
and `read_png` decodes the files this module writes.
"""
import struct
import zlib

import numpy as np

from . import abi


def unpack(packed, width, height, bgra=False):
    """(H, W, 4) RGBA uint8 from packed 32-bit pixels (byte 0 = R, or B with
    bgra)."""
    b = np.ascontiguousarray(packed).view(np.uint8).reshape(height, width, 4)
    return b[..., [2, 1, 0, 3]].copy() if bgra else b.copy()


def readback(out_rgba8, width, height, opts_flags=0):
    """Copies a device (torch) or host image of packed pixels to an (H, W, 4)
    RGBA array, honouring RTG_TRACE_OUTPUT_BGRA8 in the trace flags."""
    host = out_rgba8.cpu().numpy() if hasattr(out_rgba8, "cpu") else np.asarray(out_rgba8)
    return unpack(host.view(np.uint32), width, height, bool(opts_flags & abi.RTG_TRACE_OUTPUT_BGRA8))


def _chunk(tag, data):
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def encode_png(rgba):
    """PNG bytes of an (H, W, 4) uint8 array (colour type 6, 8 bits)."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w, c = rgba.shape
    if c != 4:
        raise ValueError("expected (H, W, 4) RGBA")
    rows = np.concatenate([np.zeros((h, 1), np.uint8), rgba.reshape(h, w * 4)], axis=1)  # filter 0 per row
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(rows.tobytes(), 6)) +
            _chunk(b"IEND", b""))


def write_png(path, rgba):
    with open(path, "wb") as f:
        f.write(encode_png(rgba))


def read_png(data):
    """(H, W, 4) array of an 8-bit RGBA PNG with filter 0 rows (what
    encode_png writes)."""
    if isinstance(data, str):
        data = open(data, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG")
    pos, idat = 8, b""
    w = h = None
    while pos < len(data):
        n, tag = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if zlib.crc32(tag + body) & 0xFFFFFFFF != struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]:
            raise ValueError(f"bad CRC in {tag!r}")
        if tag == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
            if (depth, ctype, interlace) != (8, 6, 0):
                raise ValueError("only 8-bit RGBA, non-interlaced")
        elif tag == b"IDAT":
            idat += body
        elif tag == b"IEND":
            break
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    if np.any(raw[:, 0] != 0):
        raise ValueError("only filter-0 rows")
    return raw[:, 1:].reshape(h, w, 4).copy()
