"""Bit-exact parity of the HIP draw path with the CPU oracle (GPU only)."""
import sys

import pytest

import oracle_lib
from gpu_util import compare, render_gpu
from cpvulkan_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(ctx, scene, label):
    print(f"[parity] {label}: {scene.name} {scene.width}x{scene.height} prims={scene.primitive_count}", flush=True)
    ref = oracle_lib.render_scene(scene)
    gpu = render_gpu(ctx, scene)
    compare(scene, gpu, ref, label)


@pytest.mark.parametrize("cull", ["NONE", "BACK"])
def test_cube(gpu_ctx, cull):
    _run(gpu_ctx, W.cube_scene(cull=cull), f"cube cull={cull}")


@pytest.mark.parametrize("tile,waves", [(16, 4), (8, 1), (16, 1), (16, 2), (32, 1), (32, 2), (32, 4)])
def test_soup_tile_configs(gpu_ctx, tile, waves):
    gpu_ctx.set_config(tile_size=tile, waves_per_tile=waves)
    try:
        _run(gpu_ctx, W.soup_scene(20000, 480, 270, seed=3), f"soup tile={tile} waves={waves}")
    finally:
        gpu_ctx.set_config()


def test_soup_1080p(gpu_ctx):
    _run(gpu_ctx, W.soup_scene(100000, 1920, 1080, seed=11), "soup 1080p 100k")
