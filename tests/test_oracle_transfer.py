"""Pins the transfer restatements of the oracle: FillBuffer's tail bytes
(a hand-worked case of CommandBuffer.cpp:307-326), rectangle clears
(only the rectangles change, each pixel holding the format's packed
value), and blits against exact transformations (identity with either
filter, mirroring, integer-factor NEAREST upscaling, a LINEAR 2:1
reduction within one LSB of the mean of each 2x2 block); plus the Python
format-size table against the oracle's."""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np
import pytest

import oracle_lib
from cpvulkan_amd import abi

F = abi.FORMAT
RGBA8 = F["R8G8B8A8_UNORM"]


def test_format_sizes_match_the_oracle():
    for name, code in F.items():
        if code:
            assert abi.FORMAT_BYTES[code] == oracle_lib.lib().cpvref_format_size(code), name


def test_fill_buffer_tail_bytes():
    """size 7, data 0x44332211: one word (11 22 33 44), then the three tail
    bytes 11 22 33 written from byte size / 4 = 1 on."""
    b = np.full(8, 0xEE, np.uint8)
    oracle_lib.fill_buffer(b, 7, 0x44332211)
    assert b.tolist() == [0x11, 0x11, 0x22, 0x33, 0xEE, 0xEE, 0xEE, 0xEE]
    b = np.zeros(16, np.uint8)
    oracle_lib.fill_buffer(b, 12, 0x01020304, 4)
    assert b.view(np.uint32).tolist() == [0, 0x01020304, 0x01020304, 0x01020304]


def test_clear_rects_touch_only_the_rects():
    W, H = 9, 7
    img = np.full(W * H * 4, 0x5A, np.uint8)
    value = [struct.unpack("<I", struct.pack("<f", v))[0] for v in (1.0, 0.0, 0.5, 0.2)]
    oracle_lib.clear_rects(img, RGBA8, W, H, value, [abi.Rect(2, 1, 3, 4)])
    px = img.reshape(H, W, 4)
    inside = np.zeros((H, W), bool)
    inside[1:5, 2:5] = True
    assert (px[inside] == [255, 0, 128, 51]).all()       # round-half-away of 127.5 and 51.0
    assert (px[~inside] == 0x5A).all()


def _blit(src, sw, sh, dst, dw, dh, sr, dr, filt, fmt=RGBA8):
    reg = abi.BlitRegion((C.c_int32 * 4)(*sr), (C.c_int32 * 4)(*dr), 1, 0)
    oracle_lib.blit_image(src.copy(), fmt, sw, sh, dst, fmt, dw, dh, [reg], filt)


@pytest.mark.parametrize("filt", [0, 1])
def test_blit_identity_and_mirror(filt):
    rng = np.random.Generator(np.random.PCG64(1))
    src = rng.integers(0, 256, size=6 * 5 * 4, dtype=np.uint8)
    dst = np.zeros_like(src)
    _blit(src, 6, 5, dst, 6, 5, (0, 0, 6, 5), (0, 0, 6, 5), filt)
    assert np.array_equal(dst, src)
    _blit(src, 6, 5, dst, 6, 5, (0, 0, 6, 5), (6, 5, 0, 0), filt)
    assert np.array_equal(dst.reshape(5, 6, 4), src.reshape(5, 6, 4)[::-1, ::-1])


def test_blit_nearest_upscale_repeats_texels():
    rng = np.random.Generator(np.random.PCG64(2))
    src = rng.integers(0, 256, size=4 * 3 * 4, dtype=np.uint8)
    dst = np.zeros(12 * 6 * 4, np.uint8)
    _blit(src, 4, 3, dst, 12, 6, (0, 0, 4, 3), (0, 0, 12, 6), 0)
    want = np.repeat(np.repeat(src.reshape(3, 4, 4), 2, axis=0), 3, axis=1)
    assert np.array_equal(dst.reshape(6, 12, 4), want)


def test_blit_linear_half_is_the_block_mean():
    rng = np.random.Generator(np.random.PCG64(3))
    src = rng.integers(0, 256, size=8 * 6 * 4, dtype=np.uint8)
    dst = np.zeros(4 * 3 * 4, np.uint8)
    _blit(src, 8, 6, dst, 4, 3, (0, 0, 8, 6), (0, 0, 4, 3), 1)
    mean = src.reshape(3, 2, 4, 2, 4).astype(np.float64).mean(axis=(1, 3))
    assert np.abs(dst.reshape(3, 4, 4).astype(np.float64) - mean).max() <= 1.0


def _tref():
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
    import transfer_ref
    return transfer_ref


def test_row_copies():
    T = _tref()
    rng = np.random.Generator(np.random.PCG64(8))
    buf = rng.integers(0, 256, size=4096, dtype=np.uint8)
    # a whole level is one contiguous copy
    lvl = np.zeros(16 * 8 * 4, np.uint8)
    T.buffer_to_image(buf, lvl, 16, 4, (100, 0, (0, 0), (16, 8)))
    assert np.array_equal(lvl, buf[100:100 + lvl.size])
    # buffer -> image -> buffer with the same region returns the rows and leaves
    # the bufferRowLength padding between them untouched
    lvl[:] = 0
    T.buffer_to_image(buf, lvl, 16, 4, (40, 11, (3, 2), (5, 4)))
    back = np.full(4096, 0xAB, np.uint8)
    T.image_to_buffer(lvl, 16, 4, back, (40, 11, (3, 2), (5, 4)))
    for r in range(4):
        row = slice(40 + r * 11 * 4, 40 + r * 11 * 4 + 20)
        assert np.array_equal(back[row], buf[row])
        gap = slice(40 + r * 11 * 4 + 20, 40 + (r + 1) * 11 * 4)
        if r < 3:
            assert (back[gap] == 0xAB).all()
    # only the region's pixels of the level were written
    img = lvl.reshape(8, 16, 4)
    mask = np.zeros((8, 16), bool)
    mask[2:6, 3:8] = True
    assert (img[~mask] == 0).all()


def test_d32s8_clear_keeps_padding():
    """SetPixelDepthStencil of D32_SFLOAT_S8_UINT stores the float and the
    stencil byte and leaves bytes 5-7 (ImageCompiler.cpp:895-945)."""
    img = np.full(3 * 2 * 8, 0x77, np.uint8)
    oracle_lib.clear_rects(img, F["D32_SFLOAT_S8_UINT"], 3, 2, [struct.unpack("<I", struct.pack("<f", 0.5))[0], 9, 0, 0],
                           [abi.Rect(0, 0, 3, 2)])
    px = img.reshape(6, 8)
    assert (px[:, 0:4].copy().view(np.float32) == 0.5).all()
    assert (px[:, 4] == 9).all() and (px[:, 5:] == 0x77).all()
