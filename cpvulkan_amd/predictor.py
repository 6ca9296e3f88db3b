"""Training the tile-cost predictor from traces the rasteriser records.

    traces = collect_traces(ctx, [(materialized, None), (None, jobs)], configs)
    mlp = train_predictor(ctx, traces, samples=10_000)
    ctx.set_predictor(mlp); ctx.set_config(..., scheduler=abi.SCHED["PREDICTED"])

A trace row is one rendered tile: its 12 binning features (computed on the
device after binning, before rasterisation) and the shader cycles its
wavefronts spent.  The network learns log2(cycles); the persistent raster
kernel then takes tiles in descending predicted cost (longest first), which
shortens the tail of a frame whose tiles are very uneven.  Rendering
results never depend on the model: it only orders the queue.

Schedule quality is compared with the counter-ordered queue's on the
measured per-tile costs: `simulate_makespan` list-schedules tiles in a
given order onto the resident wavefront slots and returns the finish time.
"""
from __future__ import annotations

import numpy as np

from . import abi


def _record(ctx, fn, tile, waves, count_fragments):
    cfg = ctx.config()
    saved = abi.Config.from_buffer_copy(cfg)
    cfg.tile_size, cfg.waves_per_tile, cfg.record_costs = tile, waves, 1
    cfg.count_fragments = count_fragments
    ctx.apply_config(cfg)
    try:
        fn()
        ctx.synchronize()
        return ctx.tile_traces()
    finally:
        ctx.apply_config(saved)


def collect_traces(ctx, passes=(), job_sets=(), configs=((16, 4, 0),), reset=None) -> dict:
    """Render every materialised pass (workloads.Materialized) and every job
    set (list of abi.JobDesc) once per (tile, waves, count_fragments)
    configuration with cost recording on; returns concatenated traces
    {features [n, 12], cycles [n]} of non-empty tiles."""
    feats, cyc = [], []
    for tile, waves, count in configs:
        for m in passes:
            def run(m=m):
                m.reset_attachments()
                ctx.render_pass(m.pass_desc, m.pipelines, m.draws, m.draw_pipeline)
            tr = _record(ctx, run, tile, waves, count)
            feats.append(tr["features"])
            cyc.append(tr["cycles"])
        for jobs in job_sets:
            tr = _record(ctx, lambda jobs=jobs: ctx.render_jobs(jobs), tile, waves, count)
            feats.append(tr["features"])
            cyc.append(tr["cycles"])
    f = np.concatenate(feats) if feats else np.zeros((0, 12), np.float32)
    c = np.concatenate(cyc) if cyc else np.zeros(0, np.float32)
    keep = c > 0
    return dict(features=f[keep], cycles=c[keep])


def sample_traces(traces: dict, n: int, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """n rows (without replacement when possible) as (x, log2 cycles)."""
    rng = np.random.default_rng(seed)
    rows = traces["features"].shape[0]
    idx = rng.choice(rows, size=min(n, rows), replace=False) if rows else np.zeros(0, np.int64)
    x = traces["features"][idx].astype(np.float32)
    y = np.log2(np.maximum(traces["cycles"][idx], 1.0)).astype(np.float32)
    return x, y


def train_predictor(ctx, traces: dict, samples: int = 10_000, steps: int = 3000, batch: int = 1024, lr: float = 2e-3,
                    seed: int = 0, holdout: float = 0.1):
    """Trains a fresh predictor on the device; returns (mlp, report) where
    the report holds train / held-out mean squared error of log2 cycles and
    the held-out rank correlation of predicted and measured cost."""
    from . import Mlp
    x, y = sample_traces(traces, samples, seed)
    n_hold = int(len(y) * holdout)
    xt, yt, xh, yh = x[n_hold:], y[n_hold:], x[:n_hold], y[:n_hold]
    mlp = Mlp(ctx, seed=seed)
    train_mse = mlp.train(xt, yt, steps=steps, batch=min(batch, len(yt)), lr=lr, seed=seed + 1)
    report = dict(rows=int(len(y)), train_mse=float(train_mse))
    if n_hold:
        ph = mlp.predict(xh)
        report["holdout_mse"] = float(np.mean((ph - yh) ** 2))
        report["holdout_rank_corr"] = float(np.corrcoef(np.argsort(np.argsort(ph)), np.argsort(np.argsort(yh)))[0, 1])
    return mlp, report


def simulate_makespan(costs: np.ndarray, order: np.ndarray, workers: int) -> float:
    """Finish time of list scheduling: tiles taken in `order`, each by the
    first worker to become free (the persistent queue's behaviour)."""
    import heapq
    free = [0.0] * workers
    heapq.heapify(free)
    end = 0.0
    for t in order:
        s = heapq.heappop(free)
        e = s + float(costs[t])
        end = max(end, e)
        heapq.heappush(free, e)
    return end
