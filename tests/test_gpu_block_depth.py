"""The raster chunk test's block-local depth bound at its conditions (GPU only).

A candidate triangle is dropped for a wave when its depth, bounded over its
box clipped to the wave's 8x8 block, fails the depth test against every
depth stored in the block.  The bound is taken only for triangles of the
draw last visited (whose pixel centres the lanes hold) when that draw's
depth range is [0, 1].  These scenes give the bound its cases: per-vertex
depths with steep gradients over small triangles (the bound is much
tighter than the triangle's depth range, so it drops many candidates),
several draws of one launch under different viewports (offset and
scaled, interleaved so the last visited draw keeps changing),
depth ranges other than [0, 1] (the bound must stay off), and the compare
ops that fold to a constant and those that do not.  Colour and depth must
equal the oracle's bit for bit, with hierarchical culling on (fragment
counting off), in every tile configuration with one pixel of a block per
lane (16x16 tiles with four waves, 8x8 tiles) and one without."""
import copy
import zlib

import numpy as np
import pytest

import oracle_lib
from cpvulkan_amd import abi
from cpvulkan_amd import workloads as W
from gpu_util import compare, render_gpu

pytestmark = pytest.mark.gpu

WIDTH, HEIGHT = 192, 128


def _draws(seed, viewports, n=12000):
    """One soup draw per viewport (per-vertex depth uniform in [0, 1]),
    split into chunks that alternate between the draws; enough triangles
    that tiles hold several chunks of 64 candidates (the bound needs a
    visited draw, so it starts with the second)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(3):
        for j, vp in enumerate(viewports):
            v = W.soup_triangles(n // 3, WIDTH, HEIGHT, int(rng.integers(1 << 30)), mu_px=float(np.log(6.0)), sigma=0.8)
            out.append(W.DrawData(vertices={0: v.view(np.uint8).reshape(-1)}, count=v.shape[0], viewport=vp))
    return out


VIEWPORT_SETS = {
    "one": [(0.0, 0.0, float(WIDTH), float(HEIGHT), 0.0, 1.0)],
    "offset+scaled": [(0.0, 0.0, float(WIDTH), float(HEIGHT), 0.0, 1.0), (9.0, 5.0, 150.0, 100.0, 0.0, 1.0)],
    "three": [(0.0, 0.0, float(WIDTH), float(HEIGHT), 0.0, 1.0), (9.0, 5.0, 150.0, 100.0, 0.0, 1.0), (30.0, 20.0, 100.0, 90.0, 0.0, 1.0)],
    "depth range": [(0.0, 0.0, float(WIDTH), float(HEIGHT), 0.2, 0.8), (0.0, 0.0, float(WIDTH), float(HEIGHT), 0.0, 1.0)],
}


@pytest.fixture
def ctx(gpu_ctx):
    saved = gpu_ctx.config()
    yield gpu_ctx
    gpu_ctx.apply_config(saved)


@pytest.mark.parametrize("tile,waves", [(16, 4), (8, 1), (16, 1)])
@pytest.mark.parametrize("op", ["LESS", "LESS_OR_EQUAL", "GREATER"])
@pytest.mark.parametrize("vps", list(VIEWPORT_SETS))
def test_block_depth_bound_matches_oracle(ctx, vps, op, tile, waves):
    label = f"block depth {vps} {op} {tile}x{waves}"
    print(f"[block_depth] {label}", flush=True)
    s = W.soup_scene(10, WIDTH, HEIGHT, seed=1)
    s = copy.deepcopy(s)
    s.pipeline["depth_compare"] = abi.COMPARE[op]
    if op == "GREATER":
        s.clear_depth = 0.0
    s.draws = _draws(zlib.crc32(label.encode()), VIEWPORT_SETS[vps])
    cfg = ctx.config()
    cfg.tile_size, cfg.waves_per_tile = tile, waves
    ctx.apply_config(cfg)
    ref = oracle_lib.render_scene(s)
    compare(s, render_gpu(ctx, s, count=False), (ref[0], ref[1], None), label + " (culling on)")
    compare(s, render_gpu(ctx, s), ref, label + " (counting)")
