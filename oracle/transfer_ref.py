"""CPU oracle of the byte-row copies of the transfer commands (test
infrastructure only: tests/ compare the ICD's results against it; nothing
in the product imports it).

Restates the row loops of CPVulkan/CommandBuffer.Copy.cpp for the linear,
single-layer 2D images of the draw path, on numpy uint8 arrays:
  * CopyBufferCommand::Process (:38-46): one memcpy per region;
  * CopyBufferToImageCommand::ProcessNormal (:501-544) and
    CopyImageToBufferCommand::ProcessNormal (:811-855): buffer row y of a
    region at bufferOffset + y * rowLength * pixel (rowLength = bufferRowLength
    or the region width), image row y at (y + offset.y) * stride +
    offset.x * pixel of the level, pixel * width bytes each;
  * CopyImageCommand::Process (:130-163): the same rows between two levels.
Whole-level regions take the reference's single memcpy of LevelSize, which
moves the same bytes.
"""
from __future__ import annotations

import numpy as np


def copy_buffer(src: np.ndarray, dst: np.ndarray, regions) -> None:
    for s, d, n in regions:
        dst[d:d + n] = src[s:s + n]


def _rows(level_width: int, px: int, offset, extent, buffer_offset: int, row_length: int):
    """(buffer byte, image byte, length) of each row of a buffer/image region."""
    x, y = offset
    w, h = extent
    rl = row_length or w
    stride = level_width * px
    for r in range(h):
        yield buffer_offset + r * rl * px, (y + r) * stride + x * px, w * px


def buffer_to_image(buf: np.ndarray, level: np.ndarray, level_width: int, px: int, region) -> None:
    """region = (bufferOffset, bufferRowLength, (x, y), (w, h)); level is the
    image level's bytes (a view into the image)."""
    boff, rl, offset, extent = region
    for b, i, n in _rows(level_width, px, offset, extent, boff, rl):
        level[i:i + n] = buf[b:b + n]


def image_to_buffer(level: np.ndarray, level_width: int, px: int, buf: np.ndarray, region) -> None:
    boff, rl, offset, extent = region
    for b, i, n in _rows(level_width, px, offset, extent, boff, rl):
        buf[b:b + n] = level[i:i + n]


def copy_image(src_level: np.ndarray, src_width: int, dst_level: np.ndarray, dst_width: int, px: int, src_off, dst_off, extent) -> None:
    w, h = extent
    for r in range(h):
        s = (src_off[1] + r) * src_width * px + src_off[0] * px
        d = (dst_off[1] + r) * dst_width * px + dst_off[0] * px
        dst_level[d:d + w * px] = src_level[s:s + w * px]
