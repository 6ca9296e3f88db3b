"""cpv_render_pass_noalloc (GPU only): the allocation-free variant of
cpv_render_pass, whose working buffers all live in one fixed device array
per GPU.  Its attachments and fragment counts must equal cpv_render_pass's
byte for byte, on consecutive calls that reuse the array with a smaller
input the second time; a pass that does not fit is refused before anything
is drawn, and a launch whose pairs outgrow the array draws nothing and is
reported by the next call."""
import numpy as np
import pytest

import oracle_lib
import scenes as S
import cpvulkan_amd
from cpvulkan_amd import abi
from cpvulkan_amd import workloads as W
from cpvulkan_amd.workloads import DeviceMemory, DrawData, materialize
from gpu_util import compare

pytestmark = pytest.mark.gpu


def _both(ctx, scene, mem):
    """The scene through cpv_render_pass, then through
    cpv_render_pass_noalloc on the same attachments reset to their initial
    contents: (colour, depth/stencil, fragments) of each."""
    m = materialize(scene, mem)
    ctx.render_pass(m.pass_desc, m.pipelines, m.draws, m.draw_pipeline)
    ctx.synchronize()
    plain = (m.read_color(), m.read_ds(), ctx.fragment_count())
    m.reset_attachments()
    ctx.render_pass_noalloc(m.pass_desc, m.pipelines, m.draws, m.draw_pipeline)
    ctx.synchronize()
    fixed = (m.read_color(), m.read_ds(), ctx.fragment_count())
    return plain, fixed


@pytest.fixture
def ctx():
    c = cpvulkan_amd.Context(0)
    yield c
    c.close()


def test_noalloc_matches_render_pass_on_consecutive_calls(ctx):
    mem = DeviceMemory(ctx)
    try:
        # first call: 60,000 blended triangles over LOAD attachments at 640x480;
        # second: the three-draw, two-pipeline pass at 160x120 (smaller)
        for label, scene in (("soup 60k", S.load_op(S.soup(60_000, 640, 480, seed=11), seed=11)),
                             ("multi-draw", S.multi_draw(seed=4))):
            print(f"[noalloc] {label}", flush=True)
            plain, fixed = _both(ctx, scene, mem)
            compare(scene, fixed, plain, f"{label}: noalloc vs render_pass")
            # and both equal the oracle
            compare(scene, fixed, oracle_lib.render_scene(scene), f"{label}: noalloc vs oracle")
        # two noalloc calls back to back, larger then smaller, with no sync between
        big = S.load_op(S.soup(40_000, 480, 320, seed=21), seed=21)
        small = S.soup(2_000, 96, 64, seed=22)
        mb, ms = materialize(big, mem), materialize(small, mem)
        ctx.render_pass_noalloc(mb.pass_desc, mb.pipelines, mb.draws, mb.draw_pipeline)
        ctx.render_pass_noalloc(ms.pass_desc, ms.pipelines, ms.draws, ms.draw_pipeline)
        ctx.synchronize()
        assert ctx.fragment_count() == oracle_lib.render_scene(small)[2]
        compare(big, (mb.read_color(), mb.read_ds(), None), oracle_lib.render_scene(big), "back to back, first")
        compare(small, (ms.read_color(), ms.read_ds(), None), oracle_lib.render_scene(small), "back to back, second")
    finally:
        mem.release()


def test_noalloc_scheduler_configurations(ctx):
    """The grid order and the counter queue (the learned order runs as the
    counter queue here), 8x8 and 32x32 tiles."""
    scene = S.load_op(S.soup(20_000, 320, 240, seed=31), seed=31)
    ref = oracle_lib.render_scene(scene)
    mem = DeviceMemory(ctx)
    try:
        for tile, waves, sched in ((16, 4, "GRID"), (16, 4, "COUNTER"), (8, 1, "PREDICTED"), (32, 4, "GRID")):
            ctx.set_config(tile_size=tile, waves_per_tile=waves, scheduler=abi.SCHED[sched])
            plain, fixed = _both(ctx, scene, mem)
            compare(scene, fixed, plain, f"{tile}x{waves} {sched}")
            compare(scene, fixed, ref, f"{tile}x{waves} {sched} vs oracle")
    finally:
        mem.release()


def _big_triangles(n, width, height, seed):
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
    return S.load_op(s, seed=seed)


def test_noalloc_refuses_what_does_not_fit(ctx):
    mem = DeviceMemory(ctx)
    try:
        # 600,000 triangles: their records alone (128 bytes each) outgrow the array
        huge = S.load_op(S.soup(600_000, 640, 480, seed=41), seed=41)
        mh = materialize(huge, mem)
        with pytest.raises(cpvulkan_amd.CpvError, match="allocation-free"):
            ctx.render_pass_noalloc(mh.pass_desc, mh.pipelines, mh.draws, mh.draw_pipeline)
        ctx.synchronize()
        assert np.array_equal(mh.read_color(), huge.initial_color)
        # 1,000 viewport-covering triangles at 1920x1080 in 8x8 tiles: 32.4M
        # pairs, more than the array holds; the launch draws nothing and the
        # next call reports it
        ctx.set_config(tile_size=8, waves_per_tile=1)
        over = _big_triangles(1000, 1920, 1080, seed=43)
        mo = materialize(over, mem)
        ctx.render_pass_noalloc(mo.pass_desc, mo.pipelines, mo.draws, mo.draw_pipeline)
        with pytest.raises(cpvulkan_amd.CpvError, match="allocation-free entry point had room"):
            ctx.synchronize()
        assert np.array_equal(mo.read_color(), over.initial_color)
        assert np.array_equal(mo.read_ds(), over.initial_ds)
        # the same pass through cpv_render_pass grows its arrays and draws it
        ref = oracle_lib.render_scene(over, threads=oracle_lib.default_threads())
        ctx.render_pass(mo.pass_desc, mo.pipelines, mo.draws, mo.draw_pipeline)
        ctx.synchronize()
        compare(over, (mo.read_color(), mo.read_ds(), ctx.fragment_count()), ref, "after the refusal, render_pass")
        # and the array serves the next call
        ctx.set_config()
        small = S.soup(3000, 160, 120, seed=44)
        plain, fixed = _both(ctx, small, mem)
        compare(small, fixed, plain, "after the refusal, noalloc")
    finally:
        mem.release()
