"""Back-to-back immediate render calls (GPU only).

Consecutive render calls on a context alternate between two sets of
working buffers, and a call's front half (vertex, setup, binning) runs on a
second stream as soon as the raster kernel that last used its set has
finished, under the previous call's raster kernel.  The results must be
those of the calls run one after the other: every frame bit-identical to
the oracle, other work enqueued in between (an upload of new vertex data)
seen by the next call, and an attachment drawn by one call and read as the
next call's vertex buffer read only after it is drawn."""
import copy

import numpy as np
import pytest

import oracle_lib
import scenes as S
from cpvulkan_amd import abi
from cpvulkan_amd import workloads as W
from cpvulkan_amd.workloads import DeviceMemory, DrawData, materialize
from gpu_util import compare

pytestmark = pytest.mark.gpu


@pytest.fixture
def ctx(gpu_ctx):
    saved = gpu_ctx.config()
    yield gpu_ctx
    gpu_ctx.apply_config(saved)


@pytest.mark.parametrize("sched", ["GRID", "COUNTER"])
def test_back_to_back_calls_bit_exact(ctx, sched):
    ctx.set_config(scheduler=abi.SCHED[sched])
    scenes = [S.soup(20_000, 320, 240, seed=61), S.multi_draw(seed=62), W.skewed_scene(60_000, 480, 270, seed=63),
              S.soup(5_000, 200, 150, seed=64)]
    refs = [oracle_lib.render_scene(s) for s in scenes]
    mem = DeviceMemory(ctx)
    try:
        ms = [materialize(s, mem) for s in scenes]
        # every scene twice (CLEAR: the second draw gives the same frame), no synchronisation in between
        for _ in range(2):
            for m in ms:
                ctx.render_pass(m.pass_desc, m.pipelines, m.draws, m.draw_pipeline)
        ctx.synchronize()
        assert ctx.fragment_count() == refs[-1][2]
        for s, m, r in zip(scenes, ms, refs):
            compare(s, (m.read_color(), m.read_ds(), None), r, f"{sched} back to back {s.name}")
    finally:
        mem.release()


def test_upload_between_calls_is_seen(ctx):
    """Draw, upload moved vertices, draw again with no synchronisation: the
    second call draws the moved vertices."""
    scene = S.soup(8_000, 256, 192, seed=71)
    moved = copy.deepcopy(scene)
    v = moved.draws[0].vertices[0].view(np.float32).reshape(-1, 8).copy()
    v[:, 0] = -v[:, 0]
    moved.draws[0].vertices[0] = v.reshape(-1).view(np.uint8)
    ref = oracle_lib.render_scene(moved)
    mem = DeviceMemory(ctx)
    try:
        m = materialize(scene, mem)
        ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
        mem.upload(m.draws[0].vertex_address[0], v)
        ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
        ctx.synchronize()
        compare(moved, (m.read_color(), m.read_ds(), ctx.fragment_count()), ref, "after the upload")
    finally:
        mem.release()


def _feeder(n_pixels, seed):
    """A soup drawn into an R32G32B32A32_SFLOAT target one pixel tall: its
    colour bytes are then read as the next draw's vertices (pos, colour)."""
    s = S.soup(400, n_pixels, 1, seed=seed)
    s.color_format = abi.FORMAT["R32G32B32A32_SFLOAT"]
    s.ds_format = 0
    s.pipeline = dict(s.pipeline, depth_test=0, depth_write=0)
    s.pipeline.pop("blend", None)
    s.clear_color = (0.1, 0.2, 0.3, 1.0)
    return s


def test_attachment_read_as_vertex_buffer(ctx):
    """Call 1 draws into a float colour attachment; call 2 reads that
    attachment as its vertex buffer (every 2 pixels one vertex: position
    then colour).  Call 2's vertex stage must wait for call 1's raster
    kernel."""
    feeder = _feeder(96, seed=81)                      # 96 pixels = 48 vertices = 16 triangles
    ref_feed = oracle_lib.render_scene(feeder)
    consumer = S.soup(16, 128, 96, seed=82)
    consumer.draws[0].vertices[0] = ref_feed[0].copy()
    consumer.pipeline = dict(consumer.pipeline, depth_test=0, depth_write=0)
    ref = oracle_lib.render_scene(consumer)
    assert (ref[0].reshape(-1, 4) != ref[0].reshape(-1, 4)[0]).any(), "the consumer draws something"
    mem = DeviceMemory(ctx)
    try:
        mf = materialize(feeder, mem)
        mc = materialize(consumer, mem)
        mc.draws[0].vertex_address[0] = mf.color_addr
        mc.draws[0].vertex_bytes[0] = feeder.color_bytes
        for _ in range(3):
            mc.reset_attachments()
            mem.upload(mf.color_addr, np.zeros(feeder.color_bytes, np.uint8))   # stale data a premature read would see
            ctx.render_pass(mf.pass_desc, mf.pipelines, mf.draws)
            ctx.render_pass(mc.pass_desc, mc.pipelines, mc.draws)
            ctx.synchronize()
            compare(feeder, (mf.read_color(), mf.read_ds(), None), ref_feed[:2] + (None,), "feeder")
            compare(consumer, (mc.read_color(), mc.read_ds(), None), ref[:2] + (None,), "consumer")
    finally:
        mem.release()
