"""Transfer and clear entry points of the C-ABI on the GPU against the CPU
oracle: region copies (device and host-visible memory, strided rows),
vkCmdFillBuffer (with the reference's tail-byte placement), vkCmdUpdateBuffer,
rectangle clears over colour and depth/stencil formats, and blits (scaled,
mirrored, LINEAR and NEAREST, integer and depth/stencil aspects).  Every
result must equal the oracle's bytes (numpy slicing for plain copies)."""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np
import pytest

import oracle_lib
from cpvulkan_amd import abi

pytestmark = pytest.mark.gpu
F = abi.FORMAT
RNG = np.random.Generator(np.random.PCG64(2024))


class Mem:
    """Device (or host-visible) allocations released at the end of a test."""

    def __init__(self, ctx):
        self.ctx, self.dev, self.host = ctx, [], []

    def alloc(self, data: np.ndarray, host: bool = False) -> int:
        n = max(int(data.nbytes), 16)
        addr = C.c_uint64()
        if host:
            assert self.ctx.lib.cpv_alloc_host(self.ctx.handle, n, C.byref(addr)) == 0
            self.host.append(addr.value)
            C.memmove(addr.value, data.ctypes.data, data.nbytes)
        else:
            assert self.ctx.lib.cpv_alloc(self.ctx.handle, n, C.byref(addr)) == 0
            self.dev.append(addr.value)
            self.ctx.upload(addr.value, data)
        return addr.value

    def read(self, addr: int, nbytes: int, host: bool = False) -> np.ndarray:
        self.ctx.synchronize()
        if host:
            out = np.empty(nbytes, np.uint8)
            C.memmove(out.ctypes.data, addr, nbytes)
            return out
        return self.ctx.download(addr, nbytes)

    def release(self):
        self.ctx.synchronize()
        for a in self.dev:
            self.ctx.lib.cpv_free(self.ctx.handle, a)
        for a in self.host:
            self.ctx.lib.cpv_free_host(self.ctx.handle, a)


@pytest.fixture
def mem(gpu_ctx):
    m = Mem(gpu_ctx)
    yield m
    m.release()


def rand_bytes(n):
    return RNG.integers(0, 256, size=n, dtype=np.uint8)


@pytest.mark.parametrize("src_host,dst_host", [(False, False), (True, False), (False, True)])
def test_copy_regions(gpu_ctx, mem, src_host, dst_host):
    src = rand_bytes(64 * 1024 + 13)
    dst = rand_bytes(80 * 1024)
    s = mem.alloc(src, src_host)
    d = mem.alloc(dst, dst_host)
    want = dst.copy()
    regions = [abi.CopyRegion(s + 5, d + 11, 0, 0, 1001, 1, 0),                      # one unaligned run
               abi.CopyRegion(s + 2048, d + 4096, 256, 320, 200, 37, 0),             # strided rows
               abi.CopyRegion(s + 30000, d + 50000, 64, 64, 64, 100, 0)]             # dense rows
    for r in regions:
        for k in range(r.rows):
            want[r.dst - d + k * r.dst_pitch:r.dst - d + k * r.dst_pitch + r.row_bytes] = \
                src[r.src - s + k * r.src_pitch:r.src - s + k * r.src_pitch + r.row_bytes]
    print(f"[transfer] copy regions src_host={src_host} dst_host={dst_host}", flush=True)
    gpu_ctx.copy_regions(regions)
    assert np.array_equal(mem.read(d, dst.nbytes, dst_host), want)


@pytest.mark.parametrize("offset,size", [(0, 4096), (12, 1000), (256, 4099), (8, 3)])
def test_fill_and_update(gpu_ctx, mem, offset, size):
    buf = rand_bytes(8192)
    d = mem.alloc(buf)
    want = buf.copy()
    oracle_lib.fill_buffer(want, size, 0xA1B2C3D4, offset)
    print(f"[transfer] fill offset={offset} size={size}", flush=True)
    gpu_ctx.fill_buffer(d + offset, size, 0xA1B2C3D4)
    assert np.array_equal(mem.read(d, buf.nbytes), want)
    upd = rand_bytes(777)
    gpu_ctx.update_buffer(d + 1000, upd)
    want[1000:1777] = upd
    assert np.array_equal(mem.read(d, buf.nbytes), want)


def f32words(*v):
    return [struct.unpack("<I", struct.pack("<f", x))[0] for x in v]


CLEAR_CASES = [
    ("R8G8B8A8_UNORM", f32words(0.2, 0.5, 1.5, -0.25)),
    ("B8G8R8A8_UNORM", f32words(0.2, 0.4, 0.6, 0.8)),
    ("R8G8B8A8_SNORM", f32words(-0.7, 0.3, 1.0, -1.2)),
    ("R32G32B32A32_SFLOAT", f32words(1.5, -2.25, 3e38, -0.0)),
    ("R16G16B16A16_UNORM", f32words(0.123, 0.456, 0.789, 1.0)),
    ("R8G8B8A8_UINT", [3, 300, 70000, 255]),
    ("R16_SINT", [(-40000) & 0xFFFFFFFF, 0, 0, 0]),
    ("R32G32_UINT", [0xDEADBEEF, 17, 0, 0]),
    ("D16_UNORM", f32words(0.37) + [0, 0, 0]),
    ("D32_SFLOAT", f32words(0.625) + [0, 0, 0]),
    ("X8_D24_UNORM_PACK32", f32words(0.9) + [0, 0, 0]),
    ("D24_UNORM_S8_UINT", f32words(0.25) + [77, 0, 0]),
    ("D32_SFLOAT_S8_UINT", f32words(1.0) + [200, 0, 0]),
    ("S8_UINT", [0, 5, 0, 0]),
]


@pytest.mark.parametrize("fmt,value", CLEAR_CASES, ids=[c[0] for c in CLEAR_CASES])
def test_clear_rects(gpu_ctx, mem, fmt, value):
    W, H = 67, 41
    code = F[fmt]
    bpp = abi.FORMAT_BYTES[code]
    img = rand_bytes(W * H * bpp)
    d = mem.alloc(img)
    rects = [abi.Rect(0, 0, W, 1), abi.Rect(3, 5, 20, 30), abi.Rect(W - 7, H - 9, 7, 9)]
    want = img.copy()
    oracle_lib.clear_rects(want, code, W, H, value, rects)
    print(f"[transfer] clear {fmt}", flush=True)
    gpu_ctx.clear_rects(abi.ImageLevel(d, code, W, H, 0), value, rects)
    got = mem.read(d, img.nbytes)
    assert np.array_equal(got, want), f"{fmt}: {np.count_nonzero(got != want)} bytes differ"


BLITS = [
    # (src rect, dst rect)
    ((0, 0, 37, 29), (0, 0, 37, 29)),        # same size
    ((0, 0, 37, 29), (5, 3, 80, 61)),        # upscale
    ((2, 1, 37, 29), (60, 50, 10, 4)),       # downscale, mirrored both ways
    ((37, 0, 0, 29), (0, 0, 50, 20)),        # source mirrored in x
]


@pytest.mark.parametrize("fmt,filter_", [("R8G8B8A8_UNORM", 1), ("R8G8B8A8_UNORM", 0), ("B8G8R8A8_UNORM", 1),
                                          ("R32G32B32A32_SFLOAT", 1), ("R16G16B16A16_UNORM", 0), ("R8G8B8A8_UINT", 0),
                                          ("R16G16_SINT", 0)])
def test_blit_colour(gpu_ctx, mem, fmt, filter_):
    SW, SH, DW, DH = 37, 29, 96, 70
    code = F[fmt]
    bpp = abi.FORMAT_BYTES[code]
    src = rand_bytes(SW * SH * bpp)
    if fmt == "R32G32B32A32_SFLOAT":
        src = RNG.uniform(-4, 4, size=SW * SH * 4).astype(np.float32).view(np.uint8).copy()
    dst = rand_bytes(DW * DH * bpp)
    s, d = mem.alloc(src), mem.alloc(dst)
    regions = [abi.BlitRegion((C.c_int32 * 4)(*sr), (C.c_int32 * 4)(*dr), 1, 0) for sr, dr in BLITS]
    want = dst.copy()
    oracle_lib.blit_image(src.copy(), code, SW, SH, want, code, DW, DH, regions, filter_)
    print(f"[transfer] blit {fmt} filter={filter_}", flush=True)
    gpu_ctx.blit_image(abi.ImageLevel(s, code, SW, SH, 0), abi.ImageLevel(d, code, DW, DH, 0), regions, filter_)
    got = mem.read(d, dst.nbytes)
    assert np.array_equal(got, want), f"{fmt}: {np.count_nonzero(got != want)} bytes differ"


@pytest.mark.parametrize("fmt,aspect", [("D32_SFLOAT", 2), ("D16_UNORM", 2), ("D24_UNORM_S8_UINT", 6), ("D24_UNORM_S8_UINT", 2),
                                        ("D24_UNORM_S8_UINT", 4), ("D32_SFLOAT_S8_UINT", 6)])
def test_blit_depth_stencil(gpu_ctx, mem, fmt, aspect):
    SW, SH, DW, DH = 33, 21, 50, 40
    code = F[fmt]
    bpp = abi.FORMAT_BYTES[code]
    src = rand_bytes(SW * SH * bpp)
    dst = rand_bytes(DW * DH * bpp)
    if "D32" in fmt:   # keep the float depths finite
        for img, n in ((src, SW * SH), (dst, DW * DH)):
            v = img.reshape(n, bpp)
            v[:, 0:4] = RNG.uniform(0, 1, size=n).astype(np.float32).view(np.uint8).reshape(n, 4)
    s, d = mem.alloc(src), mem.alloc(dst)
    regions = [abi.BlitRegion((C.c_int32 * 4)(0, 0, SW, SH), (C.c_int32 * 4)(2, 3, 48, 39), aspect, 0),
               abi.BlitRegion((C.c_int32 * 4)(SW, 0, 0, SH), (C.c_int32 * 4)(0, 0, 10, 10), aspect, 0)]
    want = dst.copy()
    oracle_lib.blit_image(src.copy(), code, SW, SH, want, code, DW, DH, regions, 0)
    print(f"[transfer] blit {fmt} aspect={aspect}", flush=True)
    gpu_ctx.blit_image(abi.ImageLevel(s, code, SW, SH, 0), abi.ImageLevel(d, code, DW, DH, 0), regions, 0)
    got = mem.read(d, dst.nbytes)
    assert np.array_equal(got, want), f"{fmt}: {np.count_nonzero(got != want)} bytes differ"


@pytest.mark.parametrize("load", ["CLEAR", "LOAD"])
def test_d32s8_attachment_keeps_padding(gpu_ctx, load):
    """A render pass on D32_SFLOAT_S8_UINT over random memory: CLEAR writes
    the depth and the stencil byte of every pixel and keeps the three
    padding bytes, as SetPixelDepthStencil does."""
    import copy
    from cpvulkan_amd import workloads as W
    from gpu_util import compare, render_gpu
    scene = copy.deepcopy(W.soup_scene(3000, 160, 96, seed=41))
    scene.ds_format = F["D32_SFLOAT_S8_UINT"]
    scene.pipeline["ds_format"] = scene.ds_format
    scene.ds_load = abi.LOAD_OP[load]
    ds = RNG.integers(0, 256, size=160 * 96 * 8, dtype=np.uint8)
    v = ds.reshape(-1, 8)
    v[:, 0:4] = RNG.uniform(0, 1, size=v.shape[0]).astype(np.float32).view(np.uint8).reshape(-1, 4)
    scene.initial_ds = ds
    print(f"[transfer] D32S8 pass {load}", flush=True)
    compare(scene, render_gpu(gpu_ctx, scene), oracle_lib.render_scene(scene), f"d32s8 {load}")
