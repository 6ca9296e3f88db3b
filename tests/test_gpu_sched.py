"""Scheduling and batching on the GPU: every scheduler (grid, counter queue,
predicted queue), the recorded hipGraph batch, replays, the autotuner and
the pair-capacity overflow check leave every pixel bit-identical to the CPU
oracle."""
import copy

import numpy as np
import pytest

import oracle_lib
from gpu_util import compare, compare_jobs, materialize_jobs, oracle_jobs, render_gpu
import cpvulkan_amd
from cpvulkan_amd import abi
from cpvulkan_amd import workloads as W

pytestmark = pytest.mark.gpu

SCHED = abi.SCHED


def _cfg(ctx, **kw):
    cfg = ctx.config()
    for k, v in kw.items():
        setattr(cfg, k, v)
    ctx.apply_config(cfg)


@pytest.fixture
def ctx(gpu_ctx):
    saved = gpu_ctx.config()
    yield gpu_ctx
    gpu_ctx.apply_config(saved)
    gpu_ctx.set_predictor(None)


@pytest.mark.parametrize("sched", ["COUNTER", "PREDICTED"])
@pytest.mark.parametrize("tile,waves", [(16, 4), (8, 1), (32, 2)])
def test_schedulers_bit_exact(ctx, sched, tile, waves):
    scene = W.skewed_scene(40_000, 640, 360, seed=7)
    ref = oracle_lib.render_scene(scene)
    mlp = None
    if sched == "PREDICTED":
        mlp = cpvulkan_amd.Mlp(ctx, seed=3)     # untrained: any order must give the same image
        ctx.set_predictor(mlp)
    _cfg(ctx, tile_size=tile, waves_per_tile=waves, scheduler=SCHED[sched])
    print(f"[sched] {sched} tile={tile} waves={waves}", flush=True)
    try:
        compare(scene, render_gpu(ctx, scene), ref, f"{sched} {tile}x{waves}")
        compare(scene, render_gpu(ctx, scene, count=False), ref[:2] + (None,), f"{sched} {tile}x{waves} hiz")
    finally:
        ctx.set_predictor(None)
        if mlp:
            mlp.close()


def test_predicted_without_model_falls_back(ctx):
    scene = W.soup_scene(5000, 320, 180, seed=9)
    ref = oracle_lib.render_scene(scene)
    _cfg(ctx, scheduler=SCHED["PREDICTED"])
    compare(scene, render_gpu(ctx, scene), ref, "predicted, no model")


def test_tile_traces(ctx):
    scene = W.skewed_scene(30_000, 480, 270, seed=2)
    _cfg(ctx, tile_size=16, waves_per_tile=4, record_costs=1)
    render_gpu(ctx, scene)
    tr = ctx.tile_traces()
    tiles = ((480 + 15) // 16) * ((270 + 15) // 16)
    assert tr["features"].shape == (tiles, 12)
    assert np.all(tr["cycles"] > 0)
    assert np.array_equal(tr["tile"], np.arange(tiles))
    f = tr["features"]
    assert np.all(np.isfinite(f))
    assert np.allclose(f[:, 7], 0.5) and np.allclose(f[:, 8], 1.0)
    # busier tiles cost more: rank correlation of primitive count and cycles
    rc = np.corrcoef(np.argsort(np.argsort(f[:, 0])), np.argsort(np.argsort(tr["cycles"])))[0, 1]
    print(f"[sched] traces: {tiles} tiles, rank corr(n, cycles) = {rc:.3f}", flush=True)
    assert rc > 0.5


def _jobs(n=10, seed=11):
    return W.job_scenes(n, seed=seed, sides=(64, 128, 256, 200), lo=12, hi=4000)


@pytest.mark.parametrize("sched", ["GRID", "COUNTER"])
def test_batch_replays_bit_exact(ctx, sched):
    scenes = _jobs()
    refs = oracle_jobs(scenes)
    _cfg(ctx, scheduler=SCHED[sched])
    mem, ms, jobs = materialize_jobs(ctx, scenes)
    try:
        total_ref = sum(r[2] for r in refs)
        with ctx.batch(jobs) as b:
            for rep in range(3):
                for m in ms:
                    m.reset_attachments()   # replays must not depend on what the last one left
                b.submit()
                n = b.wait(count=True)
                print(f"[batch] {sched} replay {rep}: {n} fragments, {b.timing():.3f} ms", flush=True)
                assert n == total_ref
                compare_jobs(scenes, ms, refs, f"batch {sched} replay {rep}")
    finally:
        mem.release()


def test_render_jobs_immediate(ctx):
    scenes = _jobs(6, seed=5)
    refs = oracle_jobs(scenes)
    mem, ms, jobs = materialize_jobs(ctx, scenes)
    try:
        ctx.render_jobs(jobs)
        ctx.synchronize()
        assert ctx.fragment_count() == sum(r[2] for r in refs)
        compare_jobs(scenes, ms, refs, "render_jobs")
    finally:
        mem.release()


def test_batch_outgrowing_its_capacity_is_redrawn(ctx):
    """Vertex data that grows after recording: the emit kernels drop pairs
    past the recorded capacity (never writing out of bounds), the submit
    draws nothing, and the next call records the batch again and redraws
    it."""
    scenes = _jobs(4, seed=8)
    mem, ms, jobs = materialize_jobs(ctx, scenes)
    try:
        with ctx.batch(jobs) as b:
            b.submit()
            b.wait()
            # grow every triangle 50x about its centroid: many more (tile, primitive) pairs
            grown = []
            for s, m in zip(scenes, ms):
                g = copy.deepcopy(s)
                v = g.draws[0].vertices[0].view(np.float32).reshape(-1, 3, 8).copy()
                c = v[:, :, 0:2].mean(axis=1, keepdims=True)
                v[:, :, 0:2] = c + (v[:, :, 0:2] - c) * 50.0
                g.draws[0].vertices[0] = v.reshape(-1).view(np.uint8)
                mem.upload(m.draws[0].vertex_address[0], v)
                grown.append(g)
            refs = oracle_jobs(grown)
            for m in ms:
                m.reset_attachments()
            b.submit()
            assert b.wait(count=True) == sum(r[2] for r in refs)
            compare_jobs(grown, ms, refs, "redrawn after regrowth")
    finally:
        mem.release()


def test_autotune(ctx):
    scenes = _jobs(8, seed=12)
    refs = oracle_jobs(scenes)
    mem, ms, jobs = materialize_jobs(ctx, scenes)
    try:
        table, best = ctx.autotune(jobs, repeats=2)
        print(f"[autotune] {table} -> {best.tile_size}x{best.waves_per_tile}", flush=True)
        assert len(table) == 7 and all(t[2] > 0 for t in table)
        assert (best.tile_size, best.waves_per_tile) == min(table, key=lambda t: t[2])[:2]
        ctx.apply_config(best)
        with ctx.batch(jobs) as b:
            for m in ms:
                m.reset_attachments()
            b.submit()
            b.wait()
        compare_jobs(scenes, ms, refs, "autotuned batch")
    finally:
        mem.release()


def _oracle_passes(scene, passes, between=None):
    """The oracle's attachments after `passes` renders of a scene in a row
    (loadOp LOAD: each pass draws on top of the previous one's output);
    between(host_materialized, k) runs after pass k."""
    from cpvulkan_amd.workloads import HostMemory, materialize
    hm = materialize(scene, HostMemory())
    for k in range(passes):
        oracle_lib.render(hm, threads=oracle_lib.default_threads())
        if between:
            between(hm, k)
    return hm.read_color(), hm.read_ds(), None


@pytest.mark.parametrize("sched", ["GRID", "COUNTER", "PREDICTED"])
def test_pipelined_submits_bit_exact(ctx, sched):
    """Back-to-back submits of one batch overlap a submit's vertex, setup
    and binning with the previous submit's raster kernel; with loadOp LOAD
    and blending every frame draws on the previous one, so the attachments
    after four submits equal four oracle passes in a row only if each
    raster kernel ran after the previous one and on its own frame's bins."""
    import scenes as S
    scene = S.load_op(W.skewed_scene(60_000, 640, 360, seed=17), seed=3)
    want = _oracle_passes(scene, 4)
    mlp = None
    if sched == "PREDICTED":
        mlp = cpvulkan_amd.Mlp(ctx, seed=5)
        ctx.set_predictor(mlp)
    _cfg(ctx, scheduler=SCHED[sched], count_fragments=0)
    mem, ms, jobs = materialize_jobs(ctx, [scene])
    try:
        with ctx.batch(jobs) as b:
            ms[0].reset_attachments()
            for _ in range(4):
                b.submit()
            b.wait()
            print(f"[pipelined] {sched}: 4 submits, last {b.timing():.3f} ms", flush=True)
            compare(scene, (ms[0].read_color(), ms[0].read_ds(), None), want, f"pipelined {sched}")
    finally:
        mem.release()
        ctx.set_predictor(None)
        if mlp:
            mlp.close()


def test_learned_batch_order_recorded_once(ctx):
    """A batch recorded with the learned scheduler computes its tile order
    once, from the geometry bound at recording time, and every submit
    reuses it.  The order is a schedule only: after the vertex buffer is
    overwritten with a frame whose clusters lie elsewhere, the submit must
    still draw that frame exactly."""
    a = W.skewed_scene(40_000, 640, 360, seed=21)
    b = W.skewed_scene(40_000, 640, 360, seed=22)
    ref_a, ref_b = oracle_lib.render_scene(a), oracle_lib.render_scene(b)
    mlp = cpvulkan_amd.Mlp(ctx, seed=7)
    ctx.set_predictor(mlp)
    _cfg(ctx, scheduler=SCHED["PREDICTED"], count_fragments=0)
    mem, ms, jobs = materialize_jobs(ctx, [a])
    try:
        with ctx.batch(jobs) as bt:
            bt.submit()
            bt.wait()
            compare(a, (ms[0].read_color(), ms[0].read_ds(), None), ref_a[:2] + (None,), "learned batch, recorded frame")
            mem.upload(ms[0].draws[0].vertex_address[0], b.draws[0].vertices[0])
            for _ in range(2):
                bt.submit()
            bt.wait()
            compare(b, (ms[0].read_color(), ms[0].read_ds(), None), ref_b[:2] + (None,), "learned batch, new geometry")
    finally:
        mem.release()
        ctx.set_predictor(None)
        mlp.close()


def test_pipelined_submits_order_with_other_work(ctx):
    """Work enqueued on the context between two submits (a clear of part of
    the colour attachment) is ordered between them: the second submit's
    front half waits for it, its raster kernel draws on the cleared pixels."""
    import scenes as S
    scene = S.load_op(W.soup_scene(20_000, 320, 200, seed=19), seed=4)
    value = [int(np.float32(v).view(np.uint32)) for v in (0.25, 0.5, 0.75, 1.0)]
    rect = [abi.Rect(13, 17, 150, 90)]

    def clear(hm, k):
        if k == 0:
            img = hm.mem.keep[hm.color_addr]
            oracle_lib.clear_rects(img[: scene.color_bytes], scene.color_format, scene.width, scene.height, value, rect)

    want = _oracle_passes(scene, 2, clear)
    _cfg(ctx, scheduler=SCHED["GRID"], count_fragments=0)
    mem, ms, jobs = materialize_jobs(ctx, [scene])
    try:
        with ctx.batch(jobs) as b:
            ms[0].reset_attachments()
            b.submit()
            ctx.clear_rects(abi.ImageLevel(ms[0].color_addr, scene.color_format, scene.width, scene.height, 0), value, rect)
            b.submit()
            b.wait()
            compare(scene, (ms[0].read_color(), ms[0].read_ds(), None), want, "submit, clear, submit")
    finally:
        mem.release()
