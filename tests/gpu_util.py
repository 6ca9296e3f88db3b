"""Helpers shared by the GPU parity tests: render a scene on the HIP path and
with the oracle, and compare."""
from __future__ import annotations

import numpy as np

import oracle_lib
from cpvulkan_amd.workloads import DeviceMemory, Scene, materialize


def render_gpu(ctx, scene: Scene, count: bool = True):
    """Render on the HIP path.  count=False turns exact fragment counting
    off, which enables hierarchical depth culling; the fragment count is
    then None."""
    mem = DeviceMemory(ctx)
    try:
        m = materialize(scene, mem)
        saved = ctx.config()
        if not count:
            cfg = ctx.config()
            cfg.count_fragments = 0
            ctx.apply_config(cfg)
        try:
            ctx.render_pass(m.pass_desc, m.pipelines, m.draws, m.draw_pipeline)
            ctx.synchronize()
        finally:
            ctx.apply_config(saved)
        frags = ctx.fragment_count() if count else None
        return m.read_color(), m.read_ds(), frags
    finally:
        mem.release()


def ctx_config(ctx):
    import ctypes as C
    from cpvulkan_amd import abi
    cfg = abi.Config()
    assert ctx.lib.cpv_get_config(ctx.handle, C.byref(cfg)) == 0
    return cfg


def compare(scene: Scene, gpu, ref, label: str = "") -> None:
    gc, gd, gf = gpu
    rc, rd, rf = ref
    if not np.array_equal(gc, rc):
        diff = np.nonzero(gc != rc)[0]
        px = diff // max(1, len(gc) // (scene.width * scene.height))
        raise AssertionError(f"{label} colour differs in {diff.size} bytes; first pixels "
                             f"{[(int(p % scene.width), int(p // scene.width)) for p in px[:5]]}; "
                             f"gpu {gc[diff[:8]]} ref {rc[diff[:8]]}")
    if not np.array_equal(gd, rd):
        diff = np.nonzero(gd != rd)[0]
        raise AssertionError(f"{label} depth/stencil differs in {diff.size} bytes (first at {diff[:5]})")
    if gf is not None:
        assert gf == rf, f"{label} fragment count {gf} != oracle {rf}"


def materialize_jobs(ctx, scenes):
    """Device copies of the job scenes and their C-ABI job descriptors."""
    from cpvulkan_amd.workloads import job_desc
    mem = DeviceMemory(ctx)
    ms = [materialize(s, mem) for s in scenes]
    return mem, ms, [job_desc(m) for m in ms]


def oracle_jobs(scenes):
    return [oracle_lib.render_scene(s) for s in scenes]


def compare_jobs(scenes, ms, refs, label=""):
    for s, m, ref in zip(scenes, ms, refs):
        compare(s, (m.read_color(), m.read_ds(), None), ref, f"{label} {s.name}")
