"""Textured indexed grid (MVP + Lambert, trilinear-mipmapped noise texture):
bit-exact parity of the HIP path with the CPU oracle (GPU only).

Covers the vertex-id shading of indexed draws, three-attribute vertex
fetch, uniform and texture bindings, and the device sampler for every
filter / wrap mode / LOD bias the sampler descriptor allows."""
import copy

import pytest

import oracle_lib
from gpu_util import compare, render_gpu
from cpvulkan_amd import abi
from cpvulkan_amd import workloads as W

pytestmark = pytest.mark.gpu

A = abi.ADDRESS


def _run(ctx, scene, label, count=True):
    print(f"[grid] {label}: {scene.name} {scene.width}x{scene.height} prims={scene.primitive_count}", flush=True)
    ref = oracle_lib.render_scene(scene)
    gpu = render_gpu(ctx, scene, count=count)
    compare(scene, gpu, ref, label)


def _with_sampler(scene, **kw):
    s = copy.deepcopy(scene)
    s.draws[0].textures[1].sampler.update(kw)
    s.name += "-" + "-".join(f"{k}{v}" for k, v in kw.items())
    return s


def test_grid_small(gpu_ctx):
    _run(gpu_ctx, W.grid_scene(n=64, width=480, height=270, tex_size=256), "grid 64")


@pytest.mark.parametrize("tile,waves", [(8, 1), (16, 2), (32, 4)])
def test_grid_tile_configs(gpu_ctx, tile, waves):
    gpu_ctx.set_config(tile_size=tile, waves_per_tile=waves)
    try:
        _run(gpu_ctx, W.grid_scene(n=96, width=640, height=360, tex_size=256), f"grid tile={tile} waves={waves}")
    finally:
        gpu_ctx.set_config()


def test_grid_depth_culling(gpu_ctx):
    _run(gpu_ctx, W.grid_scene(n=128, width=960, height=540, tex_size=512), "grid hiz", count=False)


@pytest.mark.parametrize("sampler", [
    dict(mag=abi.FILTER["NEAREST"]),
    dict(address_u=A["MIRRORED_REPEAT"], address_v=A["MIRRORED_REPEAT"]),
    dict(address_u=A["CLAMP_TO_EDGE"], address_v=A["MIRROR_CLAMP_TO_EDGE"]),
    dict(address_u=A["CLAMP_TO_BORDER"], address_v=A["CLAMP_TO_BORDER"], border=3),
    dict(address_u=A["CLAMP_TO_BORDER"], address_v=A["REPEAT"], border=4),
    dict(bias=1.5),                          # trilinear between levels 1 and 2
    dict(bias=2.25, mipmap=0, min=abi.FILTER["NEAREST"]),
    dict(bias=3.0, max_lod=1.0),
    dict(min_lod=2.5),
], ids=["nearest", "mirror", "clamp-mirrorclamp", "border3", "border4", "bias-trilinear", "bias-nearest",
        "maxlod", "minlod"])
def test_grid_samplers(gpu_ctx, sampler):
    base = W.grid_scene(n=48, width=320, height=180, tex_size=64, uv_repeat=6.0)
    _run(gpu_ctx, _with_sampler(base, **sampler), f"grid sampler {sampler}")


def test_grid_4k(gpu_ctx):
    """The full benchmark grid: 512x512 vertices, 522,242 triangles, 3840x2160."""
    _run(gpu_ctx, W.grid_scene(), "grid 4k")
