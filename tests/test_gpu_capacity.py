"""Pair capacity of immediate render calls (GPU only).

An immediate launch bins its (tile, primitive) pairs into arrays sized from
earlier launches, with no host round trip before its raster kernel; the next
call on the context checks the launch's 64-bit pair total.  A launch whose
pairs did not fit drew nothing, and is run again with arrays of the exact
size before anything else is enqueued.  A launch that needs 2^32 pairs or
more is refused with CPV_ERROR_OVERFLOW and leaves the attachments
untouched."""
import numpy as np
import pytest

import oracle_lib
import scenes as S
import cpvulkan_amd
from cpvulkan_amd import abi
from cpvulkan_amd import workloads as W
from cpvulkan_amd.workloads import DeviceMemory, DrawData, job_desc, materialize
from gpu_util import compare

pytestmark = pytest.mark.gpu


def big_triangles(n, width, height, seed, load=False):
    """n triangles, each covering the whole viewport (vertices far outside
    it), random depth and colour, alpha blended: every one meets every tile."""
    rng = np.random.Generator(np.random.PCG64(seed))
    v = np.zeros((n, 3, 8), np.float32)
    base = np.array([[-3.0, -3.0], [9.0, -3.0], [-3.0, 9.0]], np.float32)
    v[:, :, 0:2] = base[None] + rng.uniform(-0.5, 0.5, (n, 3, 2)).astype(np.float32)
    v[:, :, 2] = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    v[:, :, 3] = 1.0
    v[:, :, 4:8] = rng.uniform(0.2, 1.0, (n, 3, 4)).astype(np.float32)
    draw = DrawData(vertices={0: v.reshape(-1).view(np.uint8)}, count=n * 3,
                    viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0))
    s = W.Scene(f"big{n}", width, height, abi.FORMAT["R8G8B8A8_UNORM"], abi.FORMAT["D32_SFLOAT"], W.soup_pipeline(),
                [draw], clear_color=(0.0, 0.0, 0.0, 0.0), clear_depth=1.0)
    return S.load_op(s, seed=seed) if load else s


@pytest.fixture
def fresh_ctx():
    """A context of its own: its pair capacity starts from nothing."""
    ctx = cpvulkan_amd.Context(0)
    yield ctx
    ctx.close()


def test_launch_outgrowing_its_pair_arrays_is_run_again(fresh_ctx):
    # 1,500 viewport-covering triangles at 8x8 tiles: 48.6M pairs, far past the
    # first launch's 16 pairs per primitive + 1M
    scene = big_triangles(1500, 1920, 1080, seed=3)
    ref = oracle_lib.render_scene(scene, threads=oracle_lib.default_threads())
    fresh_ctx.set_config(tile_size=8, waves_per_tile=1)
    mem = DeviceMemory(fresh_ctx)
    try:
        m = materialize(scene, mem)
        fresh_ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
        fresh_ctx.synchronize()       # checks the launch, redraws it with arrays of the exact size
        frags = fresh_ctx.fragment_count()
        compare(scene, (m.read_color(), m.read_ds(), frags), ref, "redrawn after pair overflow")
        # the grown capacity now holds the launch: a second frame draws at once
        m.reset_attachments()
        fresh_ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
        fresh_ctx.synchronize()
        compare(scene, (m.read_color(), m.read_ds(), fresh_ctx.fragment_count()), ref, "second frame")
    finally:
        mem.release()


def test_more_than_2_32_pairs_refused(fresh_ctx):
    # 40,000 viewport-covering triangles at 3840x2160 with 8x8 tiles:
    # 40,000 x 129,600 = 5.18e9 pairs
    scene = big_triangles(40_000, 3840, 2160, seed=5, load=True)
    fresh_ctx.set_config(tile_size=8, waves_per_tile=1)
    mem = DeviceMemory(fresh_ctx)
    try:
        m = materialize(scene, mem)
        fresh_ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
        with pytest.raises(cpvulkan_amd.CpvError, match="more than 2\\^32 - 1"):
            fresh_ctx.synchronize()
        # nothing was drawn
        assert np.array_equal(m.read_color(), scene.initial_color)
        assert np.array_equal(m.read_ds(), scene.initial_ds)
        # a recorded batch refuses it when it is recorded
        with pytest.raises(cpvulkan_amd.CpvError, match="more than 2\\^32 - 1"):
            fresh_ctx.batch([job_desc(m)])
        # the context goes on working
        small = W.soup_scene(3000, 160, 120, seed=1)
        ref = oracle_lib.render_scene(small)
        ms = materialize(small, mem)
        fresh_ctx.render_pass(ms.pass_desc, ms.pipelines, ms.draws)
        fresh_ctx.synchronize()
        compare(small, (ms.read_color(), ms.read_ds(), fresh_ctx.fragment_count()), ref, "after the refusal")
    finally:
        mem.release()
