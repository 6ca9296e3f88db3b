"""The learned tile-cost scheduler end to end on the GPU: cost traces
recorded by the raster kernel on the soup, grid and batched-job workloads,
the network trained on the device, its held-out accuracy, and the measured
raster time of the skewed workload under the learned order against the
hardware grid order and the counter queue (the frame itself must stay
bit-identical)."""
import numpy as np
import pytest

import oracle_lib
from gpu_util import compare
from cpvulkan_amd import abi
from cpvulkan_amd import predictor as PR
from cpvulkan_amd import workloads as W
from cpvulkan_amd.workloads import DeviceMemory, job_desc, materialize

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(gpu_ctx):
    mem = DeviceMemory(gpu_ctx)
    soup = materialize(W.soup_scene(200_000, 1920, 1080, seed=21), mem)
    grid = materialize(W.grid_scene(n=192, width=1920, height=1080, tex_size=256), mem)
    jobs = [job_desc(materialize(s, mem)) for s in W.job_scenes(16, seed=22, sides=(256, 512, 1024), hi=50_000)]
    traces = PR.collect_traces(gpu_ctx, passes=[soup, grid], job_sets=[jobs],
                               configs=[(16, 4, 0), (16, 2, 0), (8, 1, 0)])
    mlp, report = PR.train_predictor(gpu_ctx, traces, samples=10_000, steps=1500)
    print(f"[predictor] {len(traces['cycles'])} traces, {report}", flush=True)
    yield mlp, report
    mlp.close()
    mem.release()


def test_predictor_accuracy(trained):
    _, report = trained
    assert report["rows"] == 10_000
    assert report["holdout_rank_corr"] > 0.8
    assert report["holdout_mse"] < 1.0     # log2 cycles: within ~2x on average


def _raster_ms(ctx, m, frames=7):
    """Median device time of the raster kernel over a few frames (HIP events
    around the launch on the context's stream)."""
    out = []
    for _ in range(frames):
        ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
        out.append(ctx.timing()[0])
    return float(np.median(out))


def test_predicted_order_on_skewed_workload(gpu_ctx, trained):
    """Measured raster-kernel time of the skewed frame under the hardware
    grid order, the counter queue and the learned heavy-first order: the
    learned order must not be slower than the grid order (5% allowance for
    noise), and all three draw the oracle's frame."""
    mlp, _ = trained
    scene = W.skewed_scene(600_000, 1920, 1080, seed=23)
    ref = oracle_lib.render_scene(scene)
    saved = gpu_ctx.config()
    mem = DeviceMemory(gpu_ctx)
    try:
        gpu_ctx.set_predictor(mlp)
        m = materialize(scene, mem)
        ms = {}
        for sched in ("GRID", "COUNTER", "PREDICTED"):
            gpu_ctx.set_config(tile_size=16, waves_per_tile=4, scheduler=abi.SCHED[sched], count_fragments=0)
            m.reset_attachments()
            gpu_ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
            gpu_ctx.synchronize()
            compare(scene, (m.read_color(), m.read_ds(), None), ref[:2] + (None,), f"{sched} order")
            ms[sched] = _raster_ms(gpu_ctx, m)
        print(f"[predictor] raster ms: grid {ms['GRID']:.3f} counter {ms['COUNTER']:.3f} learned {ms['PREDICTED']:.3f}",
              flush=True)
        assert ms["PREDICTED"] <= ms["GRID"] * 1.05
    finally:
        mem.release()
        gpu_ctx.apply_config(saved)
        gpu_ctx.set_predictor(None)
