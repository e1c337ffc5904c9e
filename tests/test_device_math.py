"""Exactness arguments that the device code relies on, checked on the CPU in
exact arithmetic: they replace a correctly rounded division or a 64-bit
remainder by cheaper steps that must give the same bits."""
from fractions import Fraction

import numpy as np


def rn32(x):
    """x (a Fraction) rounded to the nearest float32, ties to even."""
    f = np.float32(float(x))
    best = None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        key = (abs(Fraction(float(c)) - x), int(np.array(c).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    return np.float32(best[1])


def test_unorm8_residual_step_is_the_correctly_rounded_quotient():
    # unorm8 (csrc/rtg_device.hpp): q = c * RN(1/255), e = fma(-q, 255, c),
    # result fma(e, RN(1/255), q); each fma rounds once
    inv = np.float32(1.0) / np.float32(255.0)
    for c in range(256):
        q = rn32(Fraction(c) * Fraction(float(inv)))
        e = rn32(-Fraction(float(q)) * 255 + c)
        got = rn32(Fraction(float(e)) * Fraction(float(inv)) + Fraction(float(q)))
        want = np.float32(c) / np.float32(255.0)
        assert np.array(got).view(np.uint32) == np.array(want).view(np.uint32), c
    # the plain product is not enough
    prod = np.arange(256, dtype=np.float32) * inv
    assert np.count_nonzero(prod != np.arange(256, dtype=np.float32) / np.float32(255.0)) > 0


def wrap32(i, size):
    """wrap_coord's 32-bit steps (csrc/rtg_device.hpp) on an int i, |i| < 2^24."""
    m32 = 0xFFFFFFFF
    if size & (size - 1) == 0:
        m = (i & m32) & ((size - 1) & m32)
    elif i >= 0:
        m = i % size
    else:
        m = (size - 1 - ((-(i + 1)) % size)) & m32
    return m


def test_wrap_coord_32_bit_steps_equal_the_remainder():
    rng = np.random.default_rng(5)
    sizes = [1, 2, 3, 5, 7, 13, 17, 20, 24, 40, 48, 64, 255, 256, 1000, 1023, 4096, 65535, (1 << 24) + 3,
             (1 << 31) - 1, 1 << 31, (1 << 31) + 5, (1 << 32) - 1]
    edge = [0, 1, -1, 2, -2, (1 << 24) - 1, -(1 << 24) + 1]
    for size in sizes:
        ks = [k * size + d for k in (-3, -2, -1, 0, 1, 2) for d in (-1, 0, 1)]
        cand = edge + ks + list(rng.integers(-(1 << 24) + 1, 1 << 24, 2000))
        for i in cand:
            i = int(i)
            if not abs(i) < (1 << 24):
                continue
            assert wrap32(i, size) == i % size, (i, size)
