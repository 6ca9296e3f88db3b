"""Loader of the CPU oracle (oracle/build/libcpvref.so) for the tests.

The oracle is the checker: tests compare the HIP path against it; nothing in
the product imports this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from cpvulkan_amd import abi
from cpvulkan_amd.workloads import HostMemory, Materialized, Scene, materialize

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "build" / "libcpvref.so"

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            raise RuntimeError(f"{ORACLE_SO} is not built; run python -m cpvulkan_amd.build oracle")
        L = C.CDLL(str(ORACLE_SO))
        P = C.c_void_p
        L.cpvref_render_pass.argtypes = [C.POINTER(abi.PassDesc), C.POINTER(abi.PipelineDesc), C.POINTER(abi.DrawDesc),
                                         C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_uint64)]
        L.cpvref_render_pass_range.argtypes = [C.POINTER(abi.PassDesc), C.POINTER(abi.PipelineDesc),
                                               C.POINTER(abi.DrawDesc), C.POINTER(C.c_uint32), C.c_uint32,
                                               C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]
        L.cpvref_render_pass_parallel.argtypes = [C.POINTER(abi.PassDesc), C.POINTER(abi.PipelineDesc),
                                                  C.POINTER(abi.DrawDesc), C.POINTER(C.c_uint32), C.c_uint32,
                                                  C.c_int, C.POINTER(C.c_uint64)]
        L.cpvref_pack_f32.argtypes = [C.c_uint32, C.POINTER(C.c_float), P]
        L.cpvref_unpack_f32.argtypes = [C.c_uint32, P, C.POINTER(C.c_float)]
        L.cpvref_pack_u32.argtypes = [C.c_uint32, C.POINTER(C.c_uint32), P]
        L.cpvref_pack_i32.argtypes = [C.c_uint32, C.POINTER(C.c_int32), P]
        L.cpvref_pack_depth_stencil.argtypes = [C.c_uint32, C.c_float, C.c_uint32, P]
        L.cpvref_unpack_depth.argtypes = [C.c_uint32, P]
        L.cpvref_unpack_depth.restype = C.c_float
        L.cpvref_unpack_stencil.argtypes = [C.c_uint32, P]
        L.cpvref_unpack_stencil.restype = C.c_uint32
        L.cpvref_format_size.argtypes = [C.c_uint32]
        L.cpvref_format_size.restype = C.c_uint32
        L.cpvref_image_layout.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]
        L.cpvref_image_layout.restype = C.c_uint64
        L.cpvref_sample_implicit.argtypes = [C.POINTER(abi.TextureDesc), C.c_float, C.c_float, C.POINTER(C.c_float)]
        L.cpvref_sample_explicit.argtypes = [C.POINTER(abi.TextureDesc), C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float)]
        L.cpvref_fetch.argtypes = [C.POINTER(abi.TextureDesc), C.c_int32, C.c_int32, C.POINTER(C.c_float)]
        L.cpvref_fill_buffer.argtypes = [P, C.c_uint64, C.c_uint32]
        L.cpvref_fill_buffer.restype = None
        L.cpvref_clear_rects.argtypes = [C.POINTER(abi.ImageLevel), C.POINTER(C.c_uint32), C.POINTER(abi.Rect), C.c_uint32]
        L.cpvref_blit_image.argtypes = [C.POINTER(abi.ImageLevel), C.POINTER(abi.ImageLevel), C.POINTER(abi.BlitRegion),
                                        C.c_uint32, C.c_uint32]
        L.cpvref_apply_blend.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                         C.POINTER(abi.BlendAttachment), C.POINTER(C.c_float)]
        L.cpvref_apply_blend.restype = None
        L.cpvref_fragment.argtypes = [C.POINTER(abi.PipelineDesc), C.POINTER(abi.PassDesc), C.c_int32, C.c_int32,
                                      C.c_float, C.c_int, C.POINTER(C.c_float)]
        L.cpvref_glm.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.cpvref_glm.restype = None
        L.cpvref_pixel_coverage.argtypes = [C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                            C.c_int32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_int),
                                            C.POINTER(C.c_float)]
        _lib = L
    return _lib


def default_threads() -> int:
    """Host threads for the tile-parallel oracle: the cores this process may
    use, at most 16 (the host cores one MI355X of the test machines gets)."""
    import os
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def render(m: Materialized, prim_range: tuple[int, int] | None = None, threads: int = 1) -> int:
    """Run the oracle on a host-materialised scene; returns the fragment
    count.  threads > 1 runs the tile-parallel replay, which is bit-identical
    to the single-threaded one (tests/test_oracle_parallel.py checks it)."""
    L = lib()
    n = len(m.draws)
    parr = (abi.PipelineDesc * len(m.pipelines))(*m.pipelines)
    darr = (abi.DrawDesc * max(n, 1))(*m.draws)
    dp = (C.c_uint32 * max(n, 1))(*(m.draw_pipeline or [0] * n))
    frags = C.c_uint64()
    if prim_range is None and threads > 1:
        rc = L.cpvref_render_pass_parallel(C.byref(m.pass_desc), parr, darr, dp, n, threads, C.byref(frags))
    elif prim_range is None:
        rc = L.cpvref_render_pass(C.byref(m.pass_desc), parr, darr, dp, n, C.byref(frags))
    else:
        rc = L.cpvref_render_pass_range(C.byref(m.pass_desc), parr, darr, dp, n, prim_range[0], prim_range[1],
                                        C.byref(frags))
    if rc != 0:
        raise RuntimeError(f"oracle render failed ({rc})")
    return frags.value


def render_scene(scene: Scene, threads: int = 1) -> tuple[np.ndarray, np.ndarray, int]:
    """Oracle image of a scene: (colour bytes, depth/stencil bytes, fragments)."""
    m = materialize(scene, HostMemory())
    frags = render(m, threads=threads)
    return m.read_color(), m.read_ds(), frags


def fill_buffer(buf: np.ndarray, size: int, data: int, offset: int = 0) -> None:
    """FillBuffer on a host copy (uint8 array) at byte offset."""
    assert buf.dtype == np.uint8 and buf.flags.c_contiguous
    lib().cpvref_fill_buffer(buf.ctypes.data + offset, size, data & 0xFFFFFFFF)


def _level(img: np.ndarray, fmt: int, width: int, height: int) -> abi.ImageLevel:
    assert img.dtype == np.uint8 and img.flags.c_contiguous
    return abi.ImageLevel(img.ctypes.data, fmt, width, height, 0)


def clear_rects(img: np.ndarray, fmt: int, width: int, height: int, value_words, rects) -> None:
    """ClearImage / ClearAttachments on a host image (uint8 array of one level)."""
    v = (C.c_uint32 * 4)(*[int(w) & 0xFFFFFFFF for w in value_words])
    arr = (abi.Rect * max(len(rects), 1))(*rects)
    assert lib().cpvref_clear_rects(C.byref(_level(img, fmt, width, height)), v, arr, len(rects)) == 0


def blit_image(src: np.ndarray, sfmt: int, sw: int, sh: int, dst: np.ndarray, dfmt: int, dw: int, dh: int, regions,
               filter_: int) -> None:
    arr = (abi.BlitRegion * max(len(regions), 1))(*regions)
    assert lib().cpvref_blit_image(C.byref(_level(src, sfmt, sw, sh)), C.byref(_level(dst, dfmt, dw, dh)), arr, len(regions),
                                   filter_) == 0
