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
        L.cpvref_apply_blend.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                         C.POINTER(abi.BlendAttachment), C.POINTER(C.c_float)]
        L.cpvref_apply_blend.restype = None
        L.cpvref_pixel_coverage.argtypes = [C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                            C.c_int32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_int),
                                            C.POINTER(C.c_float)]
        _lib = L
    return _lib


def render(m: Materialized, prim_range: tuple[int, int] | None = None) -> int:
    """Run the oracle on a host-materialised scene; returns the fragment count."""
    L = lib()
    n = len(m.draws)
    parr = (abi.PipelineDesc * len(m.pipelines))(*m.pipelines)
    darr = (abi.DrawDesc * max(n, 1))(*m.draws)
    dp = (C.c_uint32 * max(n, 1))(*(m.draw_pipeline or [0] * n))
    frags = C.c_uint64()
    if prim_range is None:
        rc = L.cpvref_render_pass(C.byref(m.pass_desc), parr, darr, dp, n, C.byref(frags))
    else:
        rc = L.cpvref_render_pass_range(C.byref(m.pass_desc), parr, darr, dp, n, prim_range[0], prim_range[1],
                                        C.byref(frags))
    if rc != 0:
        raise RuntimeError(f"oracle render failed ({rc})")
    return frags.value


def render_scene(scene: Scene) -> tuple[np.ndarray, np.ndarray, int]:
    """Oracle image of a scene: (colour bytes, depth/stencil bytes, fragments)."""
    m = materialize(scene, HostMemory())
    frags = render(m)
    return m.read_color(), m.read_ds(), frags
