"""Parity with the oracle at the benchmark configurations' full sizes:

  soup    1920x1080, 1M alpha-blended triangles (exact fragment count, and
          the timed frame's form with hierarchical depth culling)
  skewed  3840x2160, 4M alpha-blended triangles from 8 Gaussian clusters,
          under the hardware-ordered grid, the counter-ordered persistent
          queue and the learned order (a network trained on the frame's own
          tile-cost traces)
  jobs    the 256-job Zipf batch (sides 256..4096, up to ~190k triangles a
          job) recorded as one hipGraph batch, in the autotuned tile
          configuration and in every 32x32 configuration (1, 2 and 4 waves)

These are the sizes where the binning arrays, the pair counts and the
per-tile lists are largest.  The oracle images come from the tile-parallel
oracle (bit-identical to the single-threaded replay,
tests/test_oracle_parallel.py); the job batch is compared by per-job
digests so that host memory stays bounded."""
import hashlib

import numpy as np
import pytest

import oracle_lib
from gpu_util import compare, render_gpu
import cpvulkan_amd
from cpvulkan_amd import abi
from cpvulkan_amd import predictor as PR
from cpvulkan_amd import workloads as W
from cpvulkan_amd.workloads import DeviceMemory, HostMemory, job_desc, materialize

pytestmark = pytest.mark.gpu

THREADS = oracle_lib.default_threads()


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


def test_soup_1m_1080p(ctx):
    scene = W.soup_scene(1_000_000, 1920, 1080, seed=2)
    ref = oracle_lib.render_scene(scene, threads=THREADS)
    print(f"[scale] soup 1M: oracle {ref[2]} fragments", flush=True)
    _cfg(ctx, tile_size=16, waves_per_tile=4, scheduler=abi.SCHED["GRID"])
    compare(scene, render_gpu(ctx, scene), ref, "soup 1M counted")
    compare(scene, render_gpu(ctx, scene, count=False), ref[:2] + (None,), "soup 1M hiz")


@pytest.fixture(scope="module")
def skewed():
    scene = W.skewed_scene(4_000_000, 3840, 2160, seed=5)
    ref = oracle_lib.render_scene(scene, threads=THREADS)
    print(f"[scale] skewed 4M: oracle {ref[2]} fragments", flush=True)
    return scene, ref


@pytest.mark.parametrize("sched", ["GRID", "COUNTER", "PREDICTED"])
def test_skewed_4m_4k(ctx, skewed, sched):
    scene, ref = skewed
    mlp = None
    try:
        if sched == "PREDICTED":
            # the learned order: a network trained on this frame's own traces
            mem = DeviceMemory(ctx)
            try:
                m = materialize(scene, mem)
                traces = PR.collect_traces(ctx, passes=[m], configs=[(16, 4, 0)])
            finally:
                mem.release()
            mlp, report = PR.train_predictor(ctx, traces, samples=10_000, steps=800)
            print(f"[scale] skewed predictor {report}", flush=True)
            assert report["holdout_rank_corr"] > 0.5
            ctx.set_predictor(mlp)
        _cfg(ctx, tile_size=16, waves_per_tile=4, scheduler=abi.SCHED[sched], record_costs=0)
        print(f"[scale] skewed 4M {sched}", flush=True)
        compare(scene, render_gpu(ctx, scene), ref, f"skewed {sched} counted")
        compare(scene, render_gpu(ctx, scene, count=False), ref[:2] + (None,), f"skewed {sched} hiz")
    finally:
        ctx.set_predictor(None)
        if mlp:
            mlp.close()


def _digest(colour: np.ndarray, ds: np.ndarray) -> tuple[str, str]:
    return hashlib.sha256(colour.tobytes()).hexdigest(), hashlib.sha256(ds.tobytes()).hexdigest()


@pytest.fixture(scope="module")
def jobs_256():
    scenes = W.job_scenes(256, seed=4)
    digests, total = [], 0
    for s in scenes:
        c, d, f = oracle_lib.render_scene(s, threads=THREADS)
        digests.append(_digest(c, d))
        total += f
    print(f"[scale] 256 jobs: {sum(s.primitive_count for s in scenes)} triangles, oracle {total} fragments", flush=True)
    return scenes, digests, total


def _check_jobs(scenes, ms, digests, label):
    for j, (s, m, want) in enumerate(zip(scenes, ms, digests)):
        got = _digest(m.read_color(), m.read_ds())
        if got != want:
            ref = oracle_lib.render_scene(s, threads=THREADS)
            compare(s, (m.read_color(), m.read_ds(), None), ref, f"{label} job {j} {s.name}")
            raise AssertionError(f"{label} job {j}: digest differs")


@pytest.mark.parametrize("config", ["autotuned", "32x1", "32x2", "32x4"])
def test_jobs_256_batch(ctx, jobs_256, config):
    scenes, digests, total = jobs_256
    mem = DeviceMemory(ctx)
    try:
        ms = [materialize(s, mem) for s in scenes]
        jobs = [job_desc(m) for m in ms]
        if config == "autotuned":
            table, best = ctx.autotune(jobs, repeats=1)
            print(f"[scale] autotune {table} -> {best.tile_size}x{best.waves_per_tile}", flush=True)
            ctx.apply_config(best)
        else:
            t, w = (int(v) for v in config.split("x"))
            _cfg(ctx, tile_size=t, waves_per_tile=w, scheduler=abi.SCHED["COUNTER"])
        _cfg(ctx, count_fragments=1)
        with ctx.batch(jobs) as b:
            for rep in range(2):
                for m in ms:
                    m.reset_attachments()
                b.submit()
                n = b.wait(count=True)
                print(f"[scale] jobs {config} replay {rep}: {n} fragments, {b.timing():.2f} ms", flush=True)
                assert n == total
                _check_jobs(scenes, ms, digests, f"jobs {config} replay {rep}")
    finally:
        mem.release()
