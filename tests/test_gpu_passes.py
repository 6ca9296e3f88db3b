"""Several render passes per call, recorded pass batches, and fences (GPU
only).

cpv_render_passes takes what one vkQueueSubmit replays: render passes in
order, each with its own draws (CPVulkan/Queue.cpp:11-77).  Independent
passes that share a shader program go into one launch; a pass that loads
what an earlier pass of the same call stored runs after it.  Every pass's
bytes must equal the oracle's replay of the passes one after the other.
Recorded batches of passes replay bit-exactly, and a submit whose vertex
data grew past the recorded pair capacity is recorded again and redrawn.
Fences mark the end of the work enqueued before their signal, including a
redraw that the next call makes."""
import copy

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


@pytest.fixture
def ctx(gpu_ctx):
    saved = gpu_ctx.config()
    yield gpu_ctx
    gpu_ctx.apply_config(saved)


def _independent_scenes(seed):
    """Six soups of different extents and counts, one of them LOAD with
    blending over random attachments, plus a clear-only pass."""
    out = [S.soup(n, w, h, seed=seed + k) for k, (n, w, h) in
           enumerate([(3000, 160, 120), (800, 64, 64), (12000, 320, 200), (50, 33, 17), (5000, 200, 300)])]
    out.append(S.load_op(S.soup(4000, 128, 96, seed=seed + 9), seed=seed + 9))
    empty = S.soup(10, 48, 40, seed=seed + 11)
    empty.draws = []
    out.append(empty)
    return out


def _tuple(m):
    return (m.pass_desc, m.pipelines, m.draws, m.draw_pipeline)


def test_independent_passes_take_one_launch(ctx):
    scenes = _independent_scenes(seed=100)
    refs = [oracle_lib.render_scene(s) for s in scenes]
    mem = DeviceMemory(ctx)
    try:
        ms = [materialize(s, mem) for s in scenes]
        before = ctx.launch_count()
        ctx.render_passes([_tuple(m) for m in ms])
        ctx.synchronize()
        assert ctx.launch_count() - before == 1, "seven independent passes of one shader program: one launch"
        assert ctx.fragment_count() == sum(r[2] for r in refs)
        for s, m, r in zip(scenes, ms, refs):
            compare(s, (m.read_color(), m.read_ds(), None), r, f"render_passes {s.name}")
    finally:
        mem.release()


def test_dependent_passes_run_in_order(ctx):
    """Pass 1 draws a soup; pass 2 loads pass 1's attachments and blends a
    second soup over them; pass 3, independent, has two pipelines; pass 4
    loads pass 2's result again.  The oracle replays the four in order."""
    a = S.soup(4000, 160, 120, seed=201)
    b = S.soup(4000, 160, 120, seed=202)
    c = S.multi_draw(seed=203)
    d = S.soup(1500, 160, 120, seed=204)
    ref_a = oracle_lib.render_scene(a)
    b_ref_scene = copy.deepcopy(b)
    b_ref_scene.color_load = b_ref_scene.ds_load = abi.LOAD_OP["LOAD"]
    b_ref_scene.initial_color, b_ref_scene.initial_ds = ref_a[0], ref_a[1]
    ref_b = oracle_lib.render_scene(b_ref_scene)
    d_ref_scene = copy.deepcopy(d)
    d_ref_scene.color_load = d_ref_scene.ds_load = abi.LOAD_OP["LOAD"]
    d_ref_scene.initial_color, d_ref_scene.initial_ds = ref_b[0], ref_b[1]
    ref_d = oracle_lib.render_scene(d_ref_scene)
    ref_c = oracle_lib.render_scene(c)
    mem = DeviceMemory(ctx)
    try:
        ma, mb, mc, md = (materialize(s, mem) for s in (a, b, c, d))
        # b and d draw into a's attachments, loading them
        for m in (mb, md):
            m.pass_desc.color[0] = ma.pass_desc.color[0]            # ctypes copies the structure
            m.pass_desc.depth_stencil = ma.pass_desc.depth_stencil
            m.pass_desc.color[0].load_op = abi.LOAD_OP["LOAD"]
            m.pass_desc.depth_stencil.load_op = abi.LOAD_OP["LOAD"]
        before = ctx.launch_count()
        ctx.render_passes([_tuple(ma), _tuple(mb), _tuple(mc), _tuple(md)])
        ctx.synchronize()
        launches = ctx.launch_count() - before
        assert launches == 3, f"{launches} launches: a, then b and c together, then d"
        assert ctx.fragment_count() == ref_a[2] + ref_b[2] + ref_c[2] + ref_d[2]
        compare(d, (ma.read_color(), ma.read_ds(), None), ref_d, "a, b, d on one attachment pair")
        compare(c, (mc.read_color(), mc.read_ds(), None), ref_c, "independent two-pipeline pass")
    finally:
        mem.release()


def test_pass_batch_replays(ctx):
    scenes = _independent_scenes(seed=300)
    refs = [oracle_lib.render_scene(s) for s in scenes]
    mem = DeviceMemory(ctx)
    try:
        ms = [materialize(s, mem) for s in scenes]
        with ctx.batch_passes([_tuple(m) for m in ms]) as bt:
            for rep in range(3):
                for m in ms:
                    m.reset_attachments()
                before = ctx.launch_count()
                bt.submit()
                n = bt.wait(count=True)
                assert ctx.launch_count() - before == 1
                assert n == sum(r[2] for r in refs)
                for s, m, r in zip(scenes, ms, refs):
                    compare(s, (m.read_color(), m.read_ds(), None), r, f"pass batch replay {rep} {s.name}")
    finally:
        mem.release()


def _grow(scene, factor):
    g = copy.deepcopy(scene)
    v = g.draws[0].vertices[0].view(np.float32).reshape(-1, 3, 8).copy()
    c = v[:, :, 0:2].mean(axis=1, keepdims=True)
    v[:, :, 0:2] = c + (v[:, :, 0:2] - c) * factor
    g.draws[0].vertices[0] = v.reshape(-1).view(np.uint8)
    return g, v


def test_batch_outgrowing_its_pairs_is_recorded_again(ctx):
    """Vertex data that grows after recording: the submit that needs more
    pairs draws nothing, the next call records the batch again and redraws
    it, and later submits (pipelined) use the new capacity."""
    scenes = [S.soup(2500, w, h, seed=400 + k) for k, (w, h) in enumerate([(256, 200), (128, 128), (300, 160)])]
    mem = DeviceMemory(ctx)
    try:
        ms = [materialize(s, mem) for s in scenes]
        with ctx.batch_passes([_tuple(m) for m in ms]) as bt:
            bt.submit()
            bt.wait()
            grown = []
            for s, m in zip(scenes, ms):
                g, v = _grow(s, 40.0)
                mem.upload(m.draws[0].vertex_address[0], v)
                grown.append(g)
            refs = [oracle_lib.render_scene(g) for g in grown]
            bt.submit()                    # outgrows the recorded pair arrays: draws nothing
            n = bt.wait(count=True)        # finds it, records again, redraws
            assert n == sum(r[2] for r in refs)
            for g, m, r in zip(grown, ms, refs):
                compare(g, (m.read_color(), m.read_ds(), None), r, f"redrawn {g.name}")
            for rep in range(3):           # back to back: the front half of one overlaps the raster of the previous
                bt.submit()
            n = bt.wait(count=True)
            assert n == sum(r[2] for r in refs)
            for g, m, r in zip(grown, ms, refs):
                compare(g, (m.read_color(), m.read_ds(), None), r, f"pipelined after regrowth {g.name}")
    finally:
        mem.release()


def test_fence_status_and_timeouts(ctx):
    scene = W.skewed_scene(2_000_000, 3840, 2160, seed=5)
    mem = DeviceMemory(ctx)
    try:
        m = materialize(scene, mem)
        with ctx.fence() as f, ctx.fence(signaled=True) as f_on:
            assert f.status() == abi.NOT_READY          # never signalled
            assert f.wait(0) == abi.TIMEOUT
            assert f.wait(2_000_000) == abi.TIMEOUT      # 2 ms, no work behind it
            assert f_on.status() == abi.OK and f_on.wait(0) == abi.OK
            f_on.reset()
            assert f_on.status() == abi.NOT_READY
            ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
            f.signal()
            assert f.status() == abi.NOT_READY          # a 4K frame of 2M triangles takes milliseconds
            assert f.wait(0) == abi.TIMEOUT
            assert f.wait() == abi.OK
            assert f.status() == abi.OK
            f.reset()
            assert f.status() == abi.NOT_READY
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
    return W.Scene(f"big{n}", width, height, abi.FORMAT["R8G8B8A8_UNORM"], abi.FORMAT["D32_SFLOAT"], W.soup_pipeline(),
                   [draw], clear_color=(0.0, 0.0, 0.0, 0.0), clear_depth=1.0)


def test_fence_covers_a_redraw():
    """A fence signalled right after a launch whose pairs outgrow a fresh
    context's arrays: waiting on it alone (no other call) must see the
    redrawn frame."""
    ctx = cpvulkan_amd.Context(0)
    try:
        ctx.set_config(tile_size=8, waves_per_tile=1)
        scene = _big_triangles(600, 960, 540, seed=9)
        ref = oracle_lib.render_scene(scene, threads=oracle_lib.default_threads())
        mem = DeviceMemory(ctx)
        try:
            m = materialize(scene, mem)
            with ctx.fence() as f:
                before = ctx.launch_count()
                ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
                f.signal()
                assert f.wait() == abi.OK
                # the wait itself found the overflow and enqueued the redraw
                # (cpv_launch_count resolves nothing)
                assert ctx.launch_count() - before == 2
                assert f.status() == abi.OK
                color = mem.download(m.color_addr, scene.color_bytes)
                ds = mem.download(m.ds_addr, scene.ds_bytes)
            compare(scene, (color, ds, None), ref, "read after the fence")
        finally:
            mem.release()
    finally:
        ctx.close()
