"""Programs translated from SPIR-V and compiled with hiprtc render the same
bytes as the CPU oracle running the same shaders (GPU only): the cube
(MVP vertex shader), the soup (passthrough), the textured Lambert grid and
a discarding fragment shader, plus the built-in discarding shader."""
import copy

import pytest

import oracle_lib
from gpu_util import compare, render_gpu
import cpvulkan_amd
from cpvulkan_amd import abi
from cpvulkan_amd import spirv as S
from cpvulkan_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def programs(gpu_ctx):
    progs = {}
    for vs, fs in [("vs_mvp_color", "fs_color"), ("vs_passthrough", "fs_color"), ("vs_mvp_lambert", "fs_tex_lambert"),
                   ("vs_passthrough", "fs_discard_dark")]:
        print(f"[spirv] creating program {vs} + {fs}", flush=True)
        progs[(vs, fs)] = cpvulkan_amd.Program(gpu_ctx, S.SHADERS[vs](), S.SHADERS[fs]())
    yield progs
    for p in progs.values():
        p.close()


def _run(ctx, scene, prog, label, count=True):
    """The oracle runs the scene's built-in shader kinds; the GPU runs the
    translated program (which takes precedence over the kinds)."""
    s = copy.deepcopy(scene)
    s.pipeline["program"] = prog.handle.value
    print(f"[spirv] {label}: {s.width}x{s.height} prims={s.primitive_count}", flush=True)
    ref = oracle_lib.render_scene(scene)
    compare(s, render_gpu(ctx, s, count=count), ref if count else ref[:2] + (None,), label)


def test_cube(gpu_ctx, programs):
    _run(gpu_ctx, W.cube_scene(), programs[("vs_mvp_color", "fs_color")], "cube")


def test_soup(gpu_ctx, programs):
    _run(gpu_ctx, W.soup_scene(20_000, 640, 360, seed=31), programs[("vs_passthrough", "fs_color")], "soup")


@pytest.mark.parametrize("count", [True, False])
def test_grid(gpu_ctx, programs, count):
    _run(gpu_ctx, W.grid_scene(n=96, width=640, height=360, tex_size=256), programs[("vs_mvp_lambert", "fs_tex_lambert")],
         "grid", count)


def _discard_scene():
    s = W.soup_scene(8000, 480, 270, seed=32)
    s.pipeline["fs"] = abi.FS["COLOR_DISCARD"]
    return s


def test_discard_translated(gpu_ctx, programs):
    _run(gpu_ctx, _discard_scene(), programs[("vs_passthrough", "fs_discard_dark")], "discard")


@pytest.mark.parametrize("tile,waves", [(16, 4), (8, 1), (32, 2)])
def test_discard_builtin(gpu_ctx, tile, waves):
    scene = _discard_scene()
    gpu_ctx.set_config(tile_size=tile, waves_per_tile=waves)
    try:
        compare(scene, render_gpu(gpu_ctx, scene), oracle_lib.render_scene(scene), f"builtin discard {tile}x{waves}")
    finally:
        gpu_ctx.set_config()
