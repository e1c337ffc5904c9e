"""The output step: packed RGBA8/BGRA8 pixels, PNG encoding (checked with an
independent decoder), and the oracle's BGRA8 output equal to its RGBA8 output
with red and blue swapped."""
import io

import numpy as np
import pytest

import oracle_host as oh
from rtg import abi, image, scenes


def test_unpack_rgba_and_bgra():
    px = np.array([0x04030201, 0xFF00FF10], np.uint32)  # bytes 01 02 03 04 | 10 FF 00 FF
    rgba = image.unpack(px, 2, 1)
    assert rgba.tolist() == [[[1, 2, 3, 4], [0x10, 0xFF, 0, 0xFF]]]
    assert image.unpack(px, 2, 1, bgra=True).tolist() == [[[3, 2, 1, 4], [0, 0xFF, 0x10, 0xFF]]]


def test_png_round_trip_and_independent_decoder():
    rng = np.random.default_rng(5)
    a = rng.integers(0, 256, size=(37, 53, 4), dtype=np.uint8)
    data = image.encode_png(a)
    assert np.array_equal(image.read_png(data), a)
    PIL = pytest.importorskip("PIL.Image")
    with PIL.open(io.BytesIO(data)) as im:
        assert im.mode == "RGBA" and im.size == (53, 37)
        assert np.array_equal(np.asarray(im), a)


def test_png_rejects_corruption():
    data = bytearray(image.encode_png(np.zeros((2, 2, 4), np.uint8)))
    data[20] ^= 1  # inside IHDR
    with pytest.raises(ValueError):
        image.read_png(bytes(data))


def test_oracle_bgra_output_swaps_red_and_blue():
    s = scenes.small_scene(width=24, height=16)
    hs = oh.HostScene(s)
    rgba, _, _, _ = hs.trace_image()
    o = s.trace_opts()
    o.flags = abi.RTG_TRACE_OUTPUT_BGRA8
    bgra, _, _, _ = hs.trace_image(opts=o)
    a = image.unpack(rgba, 24, 16)
    b = image.unpack(bgra, 24, 16, bgra=True)
    assert np.array_equal(a, b)
    assert not np.array_equal(rgba, bgra)
    assert (a[..., 3] == 255).all()
