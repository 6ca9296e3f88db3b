"""The learned tile-cost scheduler end to end on the GPU: cost traces
recorded by the raster kernel on the soup, grid and batched-job workloads,
the network trained on the device, its held-out accuracy, and what its
queue order gains over the counter-ordered queue on the skewed workload
(makespan of the measured per-tile costs under each order, and the frame
itself, which must stay bit-identical)."""
import numpy as np
import pytest

import oracle_lib
from gpu_util import compare, render_gpu
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


def test_predicted_order_on_skewed_workload(gpu_ctx, trained):
    mlp, _ = trained
    scene = W.skewed_scene(600_000, 1920, 1080, seed=23)
    saved = gpu_ctx.config()
    try:
        gpu_ctx.set_predictor(mlp)
        # measured per-tile costs under the counter-ordered queue
        gpu_ctx.set_config(tile_size=16, waves_per_tile=4, scheduler=abi.SCHED["COUNTER"], count_fragments=0,
                           record_costs=1)
        counter_img = render_gpu(gpu_ctx, scene, count=False)
        tr = gpu_ctx.tile_traces()
        cost = tr["cycles"] / 4.0
        pred = mlp.predict(tr["features"])
        workers = 256 * 4 * 4   # CUs x SIMDs x resident waves per SIMD
        m_index = PR.simulate_makespan(cost, np.arange(len(cost)), workers)
        m_pred = PR.simulate_makespan(cost, np.argsort(-pred, kind="stable"), workers)
        m_lpt = PR.simulate_makespan(cost, np.argsort(-cost, kind="stable"), workers)
        lb = max(cost.max(), cost.sum() / workers)
        print(f"[predictor] makespan (cycles) index {m_index:.4g} predicted {m_pred:.4g} true-cost LPT {m_lpt:.4g} "
              f"lower bound {lb:.4g}; predicted-order error vs bound {m_pred / lb - 1:.3%}", flush=True)
        assert m_pred <= m_index * 1.001
        assert m_pred <= lb * 1.05
        # the predicted queue draws the same frame
        gpu_ctx.set_config(tile_size=16, waves_per_tile=4, scheduler=abi.SCHED["PREDICTED"], count_fragments=0)
        pred_img = render_gpu(gpu_ctx, scene, count=False)
        ref = oracle_lib.render_scene(scene)
        compare(scene, pred_img, ref[:2] + (None,), "predicted queue")
        compare(scene, counter_img, ref[:2] + (None,), "counter queue")
    finally:
        gpu_ctx.apply_config(saved)
        gpu_ctx.set_predictor(None)
