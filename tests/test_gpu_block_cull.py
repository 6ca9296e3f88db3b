"""Block coverage culling at its edge cases (GPU only).

For 16x16 and 32x32 tiles the emit kernels leave out of a pair the tile
quarters (8x8 blocks, 16x16 quadrants) where some rounded edge function is
negative at all four corners of the quarter's part of the triangle's box.  The triangles here put edges through pixel
centres and along block rows and columns, make slivers a fraction of a
pixel wide on block boundaries, and put vertices on block-corner pixel
centres, under identity, offset and scaled viewports and every 16x16 and
32x32 tile configuration.  Colour, depth and the covered-fragment count must
equal the oracle's bit for bit."""
import zlib

import numpy as np
import pytest

import oracle_lib
import scenes as S
from cpvulkan_amd import workloads as W
from gpu_util import compare, render_gpu

pytestmark = pytest.mark.gpu

WIDTH, HEIGHT = 160, 120


def _ndc(px, extent):
    """NDC of pixel coordinate px (pixel centres at integers + 0.5)."""
    return (np.asarray(px, np.float64) / extent) * 2.0 - 1.0


def _vertices(xy_px: np.ndarray, rng) -> np.ndarray:
    """(n, 3, 2) pixel positions -> 3n vertices of (pos vec4, colour vec4)."""
    n = xy_px.shape[0]
    v = np.empty((n, 3, 8), np.float64)
    v[:, :, 0] = _ndc(xy_px[:, :, 0], WIDTH)
    v[:, :, 1] = _ndc(xy_px[:, :, 1], HEIGHT)
    v[:, :, 2] = rng.uniform(0.0, 1.0, size=(n, 3))
    v[:, :, 3] = 1.0
    v[:, :, 4:7] = rng.uniform(0.0, 1.0, size=(n, 3, 3))
    v[:, :, 7] = rng.uniform(0.2, 1.0, size=(n, 3))
    return v.reshape(n * 3, 8).astype(np.float32)


def _block_lines(rng, n, extent):
    """Pixel coordinates on or next to block boundaries: 8k plus a centre
    offset of 0.5 and a small step (exact centres, half pixels, slivers)."""
    k = rng.integers(0, extent // 8 + 1, size=n) * 8
    step = rng.choice([0.0, 0.5, -0.5, 1.0, -1.0, 1e-3, -1e-3, 0.25], size=n)
    return k + 0.5 + step


def snapped(seed):
    """Small triangles whose vertices sit on pixel centres near block lines."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 1500
    xy = np.empty((n, 3, 2))
    xy[:, 0, 0] = _block_lines(rng, n, WIDTH)
    xy[:, 0, 1] = _block_lines(rng, n, HEIGHT)
    for k in (1, 2):
        xy[:, k, 0] = xy[:, 0, 0] + rng.integers(-12, 13, size=n)
        xy[:, k, 1] = xy[:, 0, 1] + rng.integers(-12, 13, size=n)
    return _vertices(xy, rng)


def slivers(seed):
    """Thin triangles along block rows and columns: two vertices on one
    block line, the third a fraction of a pixel off it."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 1200
    xy = np.empty((n, 3, 2))
    along = rng.uniform(0.0, 1.0, size=(n, 2))
    horizontal = rng.random(n) < 0.5
    line_y = _block_lines(rng, n, HEIGHT)
    line_x = _block_lines(rng, n, WIDTH)
    off = rng.choice([1e-4, 1e-2, 0.3, 0.5, 0.9], size=n) * rng.choice([-1.0, 1.0], size=n)
    a = along[:, 0] * WIDTH
    b = a + rng.uniform(2.0, 40.0, size=n)
    c = along[:, 1] * HEIGHT
    d = c + rng.uniform(2.0, 40.0, size=n)
    xy[:, 0] = np.where(horizontal[:, None], np.stack([a, line_y], 1), np.stack([line_x, c], 1))
    xy[:, 1] = np.where(horizontal[:, None], np.stack([b, line_y], 1), np.stack([line_x, d], 1))
    mid_h = np.stack([(a + b) / 2, line_y + off], 1)
    mid_v = np.stack([line_x + off, (c + d) / 2], 1)
    xy[:, 2] = np.where(horizontal[:, None], mid_h, mid_v)
    return _vertices(xy, rng)


def corners(seed):
    """Triangles with a vertex on a block-corner pixel centre and edges
    leaving it at random angles."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 1500
    xy = np.empty((n, 3, 2))
    xy[:, 0, 0] = rng.integers(0, WIDTH // 8 + 1, size=n) * 8 + rng.choice([0.5, -0.5, 7.5], size=n)
    xy[:, 0, 1] = rng.integers(0, HEIGHT // 8 + 1, size=n) * 8 + rng.choice([0.5, -0.5, 7.5], size=n)
    ang = rng.uniform(0.0, 2.0 * np.pi, size=(n, 2))
    r = rng.uniform(1.0, 20.0, size=(n, 2))
    xy[:, 1:, 0] = xy[:, :1, 0] + r * np.cos(ang)
    xy[:, 1:, 1] = xy[:, :1, 1] + r * np.sin(ang)
    return _vertices(xy, rng)


CASES = {"snapped": snapped, "slivers": slivers, "corners": corners}
VIEWPORTS = [(0.0, 0.0, 160.0, 120.0, 0.0, 1.0), (5.0, 7.0, 150.0, 110.0, 0.0, 1.0), (0.0, 0.0, 200.0, 150.0, 0.0, 1.0)]


@pytest.fixture
def ctx(gpu_ctx):
    saved = gpu_ctx.config()
    yield gpu_ctx
    gpu_ctx.apply_config(saved)


@pytest.mark.parametrize("tile,waves", [(16, 4), (16, 2), (16, 1), (32, 4), (32, 2), (32, 1)])
@pytest.mark.parametrize("vp", VIEWPORTS)
@pytest.mark.parametrize("case", list(CASES))
def test_block_culling_matches_oracle(ctx, case, vp, tile, waves):
    label = f"block cull {case} vp={vp} {tile}x{waves}"
    print(f"[block_cull] {label}", flush=True)
    s = W.soup_scene(10, WIDTH, HEIGHT, seed=1)
    s.draws[0].vertices[0] = CASES[case](seed=zlib.crc32(label.encode())).view(np.uint8).reshape(-1)
    s.draws[0].count = s.draws[0].vertices[0].size // 32
    s = S.with_viewport(s, vp)
    cfg = ctx.config()
    cfg.tile_size, cfg.waves_per_tile = tile, waves
    ctx.apply_config(cfg)
    ref = oracle_lib.render_scene(s)
    compare(s, render_gpu(ctx, s), ref, label)
    compare(s, render_gpu(ctx, s, count=False), (ref[0], ref[1], None), label + " (culling on)")
