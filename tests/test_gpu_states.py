"""Parity of the HIP path with the oracle across fixed-function state,
formats, topologies and draw kinds (GPU only).  Every case compares the
colour and depth/stencil bytes and the covered-fragment count exactly."""
import numpy as np
import pytest

import oracle_lib
import scenes as S
from cpvulkan_amd import abi
from gpu_util import compare, render_gpu

pytestmark = pytest.mark.gpu

F = abi.FORMAT
CMP = abi.COMPARE
B = abi.BLEND
OP = abi.BLEND_OP
ST = abi.STENCIL


def _run(ctx, scene, label):
    print(f"[states] {label}", flush=True)
    ref = oracle_lib.render_scene(scene)
    gpu = render_gpu(ctx, scene)
    compare(scene, gpu, ref, label)
    return ref


@pytest.mark.parametrize("op", list(CMP))
@pytest.mark.parametrize("write", [0, 1])
def test_depth_compare(gpu_ctx, op, write):
    s = S.with_pipeline(S.soup(2500, 160, 120, seed=1), depth_compare=CMP[op], depth_write=write)
    s = S.with_formats(s, clear_depth=0.5)
    _run(gpu_ctx, s, f"depth {op} write={write}")


def test_depth_test_off(gpu_ctx):
    _run(gpu_ctx, S.with_pipeline(S.soup(2500, 160, 120, seed=2), depth_test=0), "depth test off")


BLENDS = {
    "add_one_one": dict(src_color=B["ONE"], dst_color=B["ONE"], color_op=OP["ADD"], src_alpha=B["ONE"], dst_alpha=B["ONE"], alpha_op=OP["ADD"]),
    "sub_colors": dict(src_color=B["SRC_COLOR"], dst_color=B["DST_COLOR"], color_op=OP["SUBTRACT"], src_alpha=B["ZERO"], dst_alpha=B["ONE"], alpha_op=OP["ADD"]),
    "revsub_const": dict(src_color=B["ONE_MINUS_DST_ALPHA"], dst_color=B["CONSTANT_COLOR"], color_op=OP["REVERSE_SUBTRACT"],
                         src_alpha=B["CONSTANT_ALPHA"], dst_alpha=B["ONE_MINUS_CONSTANT_ALPHA"], alpha_op=OP["REVERSE_SUBTRACT"]),
    "min_max": dict(src_color=B["ONE"], dst_color=B["ONE"], color_op=OP["MIN"], src_alpha=B["ONE"], dst_alpha=B["ONE"], alpha_op=OP["MAX"]),
    "saturate": dict(src_color=B["SRC_ALPHA_SATURATE"], dst_color=B["ONE_MINUS_SRC_COLOR"], color_op=OP["ADD"],
                     src_alpha=B["ONE_MINUS_DST_COLOR"], dst_alpha=B["DST_ALPHA"], alpha_op=OP["ADD"]),
    "no_blend": None,
}


@pytest.mark.parametrize("name", list(BLENDS))
def test_blend(gpu_ctx, name):
    s = S.with_pipeline(S.soup(2500, 160, 120, seed=3), blend=BLENDS[name], blend_constants=(0.25, 0.5, 0.75, 0.6),
                        depth_compare=CMP["LESS_OR_EQUAL"])
    s = S.with_formats(s, clear_color=(0.1, 0.3, 0.6, 0.8))
    _run(gpu_ctx, s, f"blend {name}")


@pytest.mark.parametrize("fmt", ["B8G8R8A8_UNORM", "R8G8B8A8_SNORM", "R16G16B16A16_UNORM", "R32G32B32A32_SFLOAT",
                                 "R8_UNORM", "R32_SFLOAT", "R8G8B8A8_UINT", "R8G8B8A8_SINT"])
def test_color_formats(gpu_ctx, fmt):
    s = S.soup(2000, 160, 120, seed=4)
    if fmt in ("R8G8B8A8_UINT", "R8G8B8A8_SINT"):
        s = S.with_pipeline(s, blend=None)
        s = S.with_formats(s, color=F[fmt], clear_color=(7, 300, 0xFFFFFFF0, 3))
    else:
        s = S.with_formats(s, color=F[fmt], clear_color=(0.3, -0.4, 0.7, 1.5))
    _run(gpu_ctx, s, f"colour {fmt}")


@pytest.mark.parametrize("fmt", ["D16_UNORM", "X8_D24_UNORM_PACK32", "D32_SFLOAT", "D24_UNORM_S8_UINT",
                                 "D32_SFLOAT_S8_UINT", "D16_UNORM_S8_UINT", "S8_UINT"])
def test_depth_formats(gpu_ctx, fmt):
    s = S.with_formats(S.soup(2000, 160, 120, seed=5), ds=F[fmt], clear_depth=0.75, clear_stencil=9)
    _run(gpu_ctx, s, f"depth format {fmt}")


STENCIL_CASES = {
    "incclamp_always": ((ST["KEEP"], ST["INC_CLAMP"], ST["DEC_CLAMP"], CMP["ALWAYS"], 0xFF, 0xFF, 1),
                        (ST["ZERO"], ST["REPLACE"], ST["INVERT"], CMP["GREATER"], 0x0F, 0xF0, 200)),
    "wrap_less": ((ST["REPLACE"], ST["INC_WRAP"], ST["DEC_WRAP"], CMP["LESS"], 0xFF, 0x7F, 3),
                  (ST["INC_CLAMP"], ST["DEC_CLAMP"], ST["KEEP"], CMP["NOT_EQUAL"], 0xFF, 0xFF, 128)),
    "equal_masked": ((ST["INVERT"], ST["REPLACE"], ST["ZERO"], CMP["EQUAL"], 0x03, 0xFF, 1),
                     (ST["DEC_WRAP"], ST["INC_WRAP"], ST["REPLACE"], CMP["LESS_OR_EQUAL"], 0xFF, 0x3C, 250)),
}


@pytest.mark.parametrize("case", list(STENCIL_CASES))
@pytest.mark.parametrize("fmt", ["D24_UNORM_S8_UINT", "D32_SFLOAT_S8_UINT", "S8_UINT"])
def test_stencil(gpu_ctx, case, fmt):
    front, back = STENCIL_CASES[case]
    s = S.with_pipeline(S.soup(3000, 160, 120, seed=6), stencil_test=1, front=front, back=back,
                        depth_compare=CMP["LESS_OR_EQUAL"])
    s = S.with_formats(s, ds=F[fmt], clear_depth=0.6, clear_stencil=126)
    _run(gpu_ctx, s, f"stencil {case} {fmt}")


def test_depth_bounds(gpu_ctx):
    s = S.with_pipeline(S.soup(2500, 160, 120, seed=7), depth_bounds_test=1, depth_bounds=(0.2, 0.7))
    _run(gpu_ctx, s, "depth bounds")


@pytest.mark.parametrize("vp", [(0.0, 0.0, 160.0, 120.0, 0.25, 0.75), (0.0, 0.0, 160.0, 120.0, 0.9, 0.1),
                                (5.0, 7.0, 150.0, 110.0, 0.0, 1.0), (0.0, 0.0, 200.0, 150.0, 0.0, 1.0)])
def test_viewport(gpu_ctx, vp):
    _run(gpu_ctx, S.with_viewport(S.soup(2500, 160, 120, seed=8), vp), f"viewport {vp}")


def test_perspective_w(gpu_ctx):
    _run(gpu_ctx, S.soup(2500, 160, 120, seed=9, w_range=(0.5, 2.5)), "perspective w")


@pytest.mark.parametrize("cull", ["NONE", "FRONT", "BACK", "FRONT_AND_BACK"])
@pytest.mark.parametrize("ff", ["CCW", "CW"])
def test_cull_front_face(gpu_ctx, cull, ff):
    s = S.with_pipeline(S.soup(2000, 160, 120, seed=10), cull=abi.CULL[cull], front_face=abi.FRONT_FACE[ff])
    _run(gpu_ctx, s, f"cull {cull} {ff}")


def test_big_and_offscreen(gpu_ctx):
    _run(gpu_ctx, S.soup(400, 160, 120, seed=11, size_px=4.0, offscreen=0.6), "big + offscreen")


def test_tiny(gpu_ctx):
    _run(gpu_ctx, S.soup(6000, 160, 120, seed=12, size_px=0.0), "sub-pixel triangles")


@pytest.mark.parametrize("fan", [False, True])
def test_strip_fan(gpu_ctx, fan):
    _run(gpu_ctx, S.strip(600, 160, 120, seed=13, fan=fan), f"{'fan' if fan else 'strip'}")


@pytest.mark.parametrize("dtype,offset", [(np.uint32, 0), (np.uint16, 0), (np.uint32, 17), (np.uint16, 5)])
def test_indexed(gpu_ctx, dtype, offset):
    s = S.indexed(S.soup(2000, 160, 120, seed=14), index_dtype=dtype, vertex_offset=offset, seed=3)
    _run(gpu_ctx, s, f"indexed {np.dtype(dtype).name} offset={offset}")


def test_instanced(gpu_ctx):
    _run(gpu_ctx, S.instanced(600, 160, 120, seed=15), "instanced")


def test_unorm_attribute(gpu_ctx):
    _run(gpu_ctx, S.unorm_colour_attribute(800, seed=16), "unorm colour attribute")


@pytest.mark.parametrize("dont_care", [False, True])
def test_load_op(gpu_ctx, dont_care):
    _run(gpu_ctx, S.load_op(S.soup(2000, 160, 120, seed=17), seed=4, dont_care=dont_care), "load op")


def test_multi_draw_pipelines(gpu_ctx):
    _run(gpu_ctx, S.multi_draw(seed=18), "three draws, two pipelines")


def test_odd_extent(gpu_ctx):
    # framebuffer not a multiple of any tile size; 3-byte depth pixels
    s = S.with_formats(S.soup(1500, 133, 77, seed=19), ds=F["D16_UNORM_S8_UINT"])
    s = S.with_viewport(s, (0.0, 0.0, 133.0, 77.0, 0.0, 1.0))
    for tile, waves in [(8, 1), (16, 4), (32, 2)]:
        gpu_ctx.set_config(tile_size=tile, waves_per_tile=waves)
        try:
            _run(gpu_ctx, s, f"odd extent tile {tile}x{waves}")
        finally:
            gpu_ctx.set_config()


def test_empty_pass_clears(gpu_ctx):
    s = S.soup(10, 96, 64, seed=20)
    s.draws = []
    ref = oracle_lib.render_scene(s)
    gpu = render_gpu(gpu_ctx, s)
    compare(s, gpu, ref, "clear only")


# ---- hierarchical depth culling (count_fragments off): images stay bit-exact

def _run_hiz(ctx, scene, label):
    print(f"[hiz] {label}", flush=True)
    ref = oracle_lib.render_scene(scene)
    gpu = render_gpu(ctx, scene, count=False)
    compare(scene, gpu, ref, label + " (hierarchical depth culling)")


@pytest.mark.parametrize("op", list(CMP))
def test_hiz_compare_ops(gpu_ctx, op):
    s = S.with_pipeline(S.soup(4000, 160, 120, seed=21), depth_compare=CMP[op])
    _run_hiz(gpu_ctx, S.with_formats(s, clear_depth=0.5), f"op {op}")


@pytest.mark.parametrize("fmt", ["D16_UNORM", "X8_D24_UNORM_PACK32", "D32_SFLOAT"])
def test_hiz_formats(gpu_ctx, fmt):
    _run_hiz(gpu_ctx, S.with_formats(S.soup(4000, 160, 120, seed=22), ds=F[fmt]), f"format {fmt}")


@pytest.mark.parametrize("vp", [(0.0, 0.0, 160.0, 120.0, 0.25, 0.75), (0.0, 0.0, 160.0, 120.0, 0.9, 0.1)])
def test_hiz_viewport_depth(gpu_ctx, vp):
    _run_hiz(gpu_ctx, S.with_viewport(S.soup(4000, 160, 120, seed=23), vp), f"viewport {vp}")


def test_hiz_perspective_and_sizes(gpu_ctx):
    _run_hiz(gpu_ctx, S.soup(4000, 160, 120, seed=24, w_range=(0.5, 2.5)), "perspective w")
    _run_hiz(gpu_ctx, S.soup(600, 160, 120, seed=25, size_px=4.0, offscreen=0.6), "big + offscreen")
    _run_hiz(gpu_ctx, S.soup(8000, 160, 120, seed=26, size_px=0.0), "tiny")


def test_hiz_multi_draw_and_load(gpu_ctx):
    _run_hiz(gpu_ctx, S.multi_draw(seed=27), "multi draw")
    _run_hiz(gpu_ctx, S.load_op(S.soup(3000, 160, 120, seed=28), seed=2), "load op")


def test_hiz_deep_1080p(gpu_ctx):
    from cpvulkan_amd import workloads as W
    _run_hiz(gpu_ctx, W.soup_scene(200000, 1920, 1080, seed=29), "soup 1080p 200k")
