"""Draw-path benchmark: frames/s of one workload on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload soup|cube|grid|skewed|jobs|transfers]
                    [--cpu-baseline full|sample|none] [--no-check]

A step is one frame: the render pass of the workload (clear, vertex stage,
setup, binning, raster/fragment) over inputs already resident in HBM.  The
default workload is the 1920x1080, 1M-triangle alpha-blended soup
(R8G8B8A8_UNORM + D32_SFLOAT, LESS, SRC_ALPHA/ONE_MINUS_SRC_ALPHA).

  skewed  3840x2160, 4M triangles from 8 Gaussian clusters, alpha blended:
          the plain single-draw path (hardware-ordered tile grid, immediate
          render call) against the batched renderer (the frame recorded as
          hipGraphs, tiles dispatched costliest-first in the order of the
          BF16 tile-cost network trained on the device on traces of the
          soup, grid and job workloads, the costliest at raised wave
          priority, consecutive submits pipelined) and against the
          counter-ordered persistent queue.  `value` is the batched
          renderer's frames/s.  The default soup line carries the same
          comparison under "skewed".
  jobs    256 independent render jobs (Zipf triangle counts, 256..4096
          square targets) recorded as one hipGraph batch after the
          autotuner picked the tile configuration; a step is one replay of
          all jobs.
  transfers  the transfer and clear commands on 1 GiB buffers / images:
          vkCmdFillBuffer, a buffer copy, a full R8G8B8A8 clear of a
          16384x16384 image, a LINEAR blit of an 8192x8192 image; `value`
          (GB/s) is the blit's algorithmic bytes (source read once,
          destination written once) over its time.  A step is one of each.

Before timing, the GPU frames are compared byte for byte with the CPU
oracle's.  The CPU baseline is the oracle timed on the host: the
tile-parallel replay over all cores (bit-identical to the single-threaded
one, which is timed too and reported beside it).  The dominant kernel's
duration comes from HIP events on the stream it runs on.  Prints one JSON
line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def fmt_name(code: int) -> str:
    from cpvulkan_amd import abi
    return next((k for k, v in abi.FORMAT.items() if v == code), str(code))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def make_scene(name: str):
    from cpvulkan_amd import workloads as W
    if name == "soup":
        return W.soup_scene(1_000_000, 1920, 1080, seed=2), "soup-1080p-1M-blend"
    if name == "cube":
        return W.cube_scene(), "cube-500x500"
    if name == "grid":
        return W.grid_scene(512, 3840, 2160, seed=3), "grid-4k-512x512-textured"
    if name == "skewed":
        return W.skewed_scene(4_000_000, 3840, 2160, seed=5), "skewed-4k-4M-blend"
    raise SystemExit(f"unknown workload {name}")


def geometry_bytes(scene) -> int:
    """Vertex, index, uniform and sampled-texture bytes a frame reads once."""
    n = 0
    for d in scene.draws:
        for v in d.vertices.values():
            n += v.nbytes
        if d.indices is not None:
            n += d.indices.nbytes
        for u in d.uniforms.values():
            n += u.nbytes
        for t in d.textures.values():
            n += t.width * t.height * 4      # level 0: the only level implicit LOD reads
    return n


def oracle_threads() -> int:
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib
    return oracle_lib.default_threads()


def oracle_frames(scenes, mode: str, threads: int = 1):
    """Oracle render of each scene: (results or None per scene, seconds).
    mode "sample" renders the first 50k triangles of each scene and scales
    the time by the triangle count (no image to compare)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib
    from cpvulkan_amd.workloads import HostMemory, materialize
    out, t_total = [], 0.0
    for s in scenes:
        hm = materialize(s, HostMemory())
        if mode == "sample":
            n = min(s.primitive_count, 50_000)
            t0 = time.perf_counter()
            oracle_lib.render(hm, (0, n))
            t_total += (time.perf_counter() - t0) * s.primitive_count / max(n, 1)
            out.append(None)
        else:
            t0 = time.perf_counter()
            frags = oracle_lib.render(hm, threads=threads)
            t_total += time.perf_counter() - t0
            out.append((hm.read_color(), hm.read_ds(), frags))
    return out, t_total


def same(a, b) -> bool:
    import numpy as np
    return (np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            and (a[2] is None or b[2] is None or a[2] == b[2]))


def measured_traffic(workload: str, tile: int, waves: int):
    """HBM bytes per raster launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (profiles/traffic.json), when they cover this
    workload and tile configuration; else None."""
    try:
        t = json.loads((ROOT / "profiles" / "traffic.json").read_text()).get(workload)
    except (OSError, ValueError):
        return None
    if not t or t.get("tile") != tile or t.get("waves_per_tile") != waves:
        return None
    return int(t["bytes_per_launch"])


def roofline(kernel_ms: float, alg_bytes: int, traffic=None) -> dict:
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 5), traffic=traffic)


def cpu_entry(dt: float, mode: str, what: str, threads: int = 1, single_dt: float | None = None) -> dict:
    """The CPU baseline: the oracle's frames/s on `threads` host cores (the
    tile-parallel replay when threads > 1), with the single-threaded
    replay's, the reference's own loop structure, beside it."""
    how = "tile-parallel over %d threads" % threads if threads > 1 else "single-threaded"
    note = f"{what} {how}, {dt:.2f} s" if mode == "full" else \
        f"{what}: first 50000 triangles per frame single-threaded, time scaled by triangle count"
    out = dict(value=round(1.0 / dt, 5), unit="frames/s", cores=threads if mode == "full" else 1, kind="port",
               sample=f"{note}; {cpu_model()}")
    if single_dt is not None:
        out["single_thread"] = dict(value=round(1.0 / single_dt, 5), unit="frames/s", cores=1,
                                    sample=f"{what} single-threaded, {single_dt:.2f} s")
    return out


def percentiles(ms: list) -> dict:
    import numpy as np
    a = np.asarray(ms, dtype=np.float64)
    return dict(median=round(float(np.median(a)), 4), p10=round(float(np.percentile(a, 10)), 4),
                p90=round(float(np.percentile(a, 90)), 4))


def bench_transfers(args, ctx):
    import ctypes as C
    import numpy as np
    from cpvulkan_amd import abi
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib
    rgba8 = abi.FORMAT["R8G8B8A8_UNORM"]
    value = [int(np.float32(v).view(np.uint32)) for v in (0.3, 0.6, 0.9, 1.0)]
    # correctness gate: a scaled, offset blit, a rectangle clear and an
    # unaligned fill against the oracle
    rng = np.random.Generator(np.random.PCG64(9))
    SW, SH, DW, DH = 301, 203, 517, 389
    src = rng.integers(0, 256, SW * SH * 4, dtype=np.uint8)
    dst = rng.integers(0, 256, DW * DH * 4, dtype=np.uint8)
    sa, da = ctx.alloc(src.nbytes), ctx.alloc(dst.nbytes)
    ctx.upload(sa, src)
    ctx.upload(da, dst)
    reg = [abi.BlitRegion((C.c_int32 * 4)(0, 0, SW, SH), (C.c_int32 * 4)(3, 2, DW - 5, DH - 1), 1, 0)]
    rect = [abi.Rect(7, 9, 200, 150)]
    ctx.blit_image(abi.ImageLevel(sa, rgba8, SW, SH, 0), abi.ImageLevel(da, rgba8, DW, DH, 0), reg, 1)
    ctx.clear_rects(abi.ImageLevel(da, rgba8, DW, DH, 0), value, rect)
    ctx.fill_buffer(sa + 4, 10001, 0xDEADBEEF)
    ctx.synchronize()
    got_d, got_s = ctx.download(da, dst.nbytes), ctx.download(sa, src.nbytes)
    want_d, want_s = dst.copy(), src.copy()
    oracle_lib.blit_image(src.copy(), rgba8, SW, SH, want_d, rgba8, DW, DH, reg, 1)
    oracle_lib.clear_rects(want_d, rgba8, DW, DH, value, rect)
    oracle_lib.fill_buffer(want_s, 10001, 0xDEADBEEF, 4)
    check = "bit-exact" if np.array_equal(got_d, want_d) and np.array_equal(got_s, want_s) else "MISMATCH"
    ctx.free(sa)
    ctx.free(da)
    log(f"transfer check {check}")
    if check != "bit-exact":
        return None, dict(error="GPU transfers differ from the CPU oracle")
    # timed, on 1 GiB buffers (launch overhead well under 1% of a command)
    N = 1 << 30
    a, b = ctx.alloc(N), ctx.alloc(N)
    big = abi.ImageLevel(a, rgba8, 16384, 16384, 0)
    bs, bd = abi.ImageLevel(a, rgba8, 8192, 8192, 0), abi.ImageLevel(b, rgba8, 8192, 8192, 0)
    blit_reg = [abi.BlitRegion((C.c_int32 * 4)(0, 0, 8192, 8192), (C.c_int32 * 4)(8192, 0, 0, 8192), 1, 0)]
    ops = dict(
        fill=(lambda: ctx.fill_buffer(a, N, 0x01020304), N),
        copy=(lambda: ctx.copy_regions([abi.CopyRegion(a, b, 0, 0, N, 1, 0)]), 2 * N),
        clear=(lambda: ctx.clear_rects(big, value, [abi.Rect(0, 0, 16384, 16384)]), N),
        blit=(lambda: ctx.blit_image(bs, bd, blit_reg, 1), 2 * 8192 * 8192 * 4),
    )
    rates, times = {}, {}
    for name, (fn, nbytes) in ops.items():
        for _ in range(args.warmup):
            fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        times[name] = dt
        rates[name] = round(nbytes / dt / 1e9, 1)
    ctx.free(a)
    ctx.free(b)
    # CPU baseline: the oracle's LINEAR blit of a 1024x1024 image, per byte
    cs = rng.integers(0, 256, 1024 * 1024 * 4, dtype=np.uint8)
    cd = np.zeros_like(cs)
    creg = [abi.BlitRegion((C.c_int32 * 4)(0, 0, 1024, 1024), (C.c_int32 * 4)(1024, 0, 0, 1024), 1, 0)]
    t0 = time.perf_counter()
    oracle_lib.blit_image(cs, rgba8, 1024, 1024, cd, rgba8, 1024, 1024, creg, 1)
    cpu_dt = time.perf_counter() - t0
    cpu = dict(value=round(2 * cs.nbytes / cpu_dt / 1e9, 4), unit="GB/s", cores=1, kind="port",
               sample=f"LINEAR blit of one 1024x1024 R8G8B8A8 image single-threaded, {cpu_dt:.2f} s; {cpu_model()}")
    blit_ms = times["blit"] * 1e3
    out = dict(
        metric="GB/s", unit="GB/s", value=rates["blit"], ms_per_step=round(sum(times.values()) * 1e3, 4),
        config=dict(workload="transfers-1GiB", fill_bytes=N, copy_bytes=N, clear="16384x16384 R8G8B8A8_UNORM",
                    blit="8192x8192 R8G8B8A8_UNORM LINEAR, mirrored in x"),
        ops_gbs=rates, ops_ms={k: round(v * 1e3, 4) for k, v in times.items()},
        roofline=roofline(blit_ms, 2 * 8192 * 8192 * 4), cpu_baseline=cpu, check=check)
    return out, None


def bench_single(args, ctx, scene, workload):
    import numpy as np
    from cpvulkan_amd.workloads import DeviceMemory, job_desc, materialize
    mem = DeviceMemory(ctx)
    m = materialize(scene, mem)
    # one frame with exact fragment counting (the count is a statistic; it
    # forces coverage evaluation of every pixel), then one frame the way the
    # timed frames run (counting off, hierarchical depth culling on); both
    # go through the correctness gate
    ctx.set_config(tile_size=args.tile, waves_per_tile=args.waves)
    ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
    ctx.synchronize()
    fragments = ctx.fragment_count()
    counted = (m.read_color(), m.read_ds(), fragments)
    ctx.set_config(tile_size=args.tile, waves_per_tile=args.waves, count_fragments=0)
    m.reset_attachments()
    ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
    ctx.synchronize()
    fast = (m.read_color(), m.read_ds(), None)
    log(f"gpu frames done: {fragments} fragments")
    cpu, check = None, "skipped"
    if args.cpu_baseline != "none" or not args.no_check:
        mode = "sample" if args.cpu_baseline == "sample" else "full"
        th = oracle_threads()
        refs, dt = oracle_frames([scene], mode, threads=th)
        single = oracle_frames([scene], "full", threads=1)[1] if mode == "full" and args.cpu_baseline == "full" else None
        cpu = cpu_entry(dt, mode, f"one frame ({scene.primitive_count} triangles)", threads=th, single_dt=single)
        if refs[0] is not None and not args.no_check:
            ok = same(counted, refs[0]) and same(fast, refs[0])
            check = "bit-exact" if ok else "MISMATCH"
            log(f"oracle frame {dt:.2f} s, check {check}")
            if not ok:
                return None, dict(error="GPU frame differs from the CPU oracle")
    # the timed frames: immediate render calls back to back, synchronised
    # before and after (the K steps of the contract)
    fps = timed(ctx, lambda: ctx.render_pass(m.pass_desc, m.pipelines, m.draws), args.steps, args.warmup)
    wall = args.steps / fps
    # device times per frame, from the same frames drawn one at a time
    # (cpv_last_timing waits for each)
    raster_ms, total_ms = [], []
    for _ in range(args.steps):
        ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
        r, t = ctx.timing()
        raster_ms.append(r)
        total_ms.append(t)
    raster_avg = float(np.mean(raster_ms))
    # the same frame recorded once (a reusable command buffer) and submitted
    # back to back: each submit's vertex, setup and binning stages run under
    # the previous submit's raster kernel
    with ctx.batch([job_desc(m)]) as rec:
        rec.submit()
        rec.wait()
        batch_fps = timed(ctx, rec.submit, args.steps, args.warmup)
        rec.wait()
    if not args.no_check:
        ok = same(fast, (m.read_color(), m.read_ds(), None))
        check = check if ok else "MISMATCH (recorded batch)"
    b_fb = scene.color_bytes + scene.ds_bytes
    b_alg = b_fb + geometry_bytes(scene)
    out = dict(
        value=round(fps, 3), ms_per_step=round(wall * 1000.0 / args.steps, 4),
        config=dict(workload=workload, width=scene.width, height=scene.height, triangles=scene.primitive_count,
                    tile=args.tile, waves_per_tile=args.waves, color=fmt_name(scene.color_format),
                    depth=fmt_name(scene.ds_format)),
        recorded_batch_fps=round(batch_fps, 3),
        fragments_per_frame=fragments, fragments_per_s=round(fragments * fps, 1),
        raster_ms=round(raster_avg, 4), frame_gpu_ms=round(float(np.mean(total_ms)), 4),
        frame_gpu_ms_pct=percentiles(total_ms), raster_ms_pct=percentiles(raster_ms),
        bytes_per_frame=dict(framebuffer=b_fb, geometry=geometry_bytes(scene)),
        roofline=roofline(raster_avg, b_fb, measured_traffic(workload, args.tile, args.waves)),
        # the whole frame's algorithmic bytes (framebuffer + geometry) at the measured frame rate
        frame_roofline=dict(bytes=b_alg, achieved=round(b_alg * fps / 1e9, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                            frac=round(b_alg * fps / 1e9 / HBM_PEAK_GBS, 5)),
        cpu_baseline=cpu, check=check,
    )
    mem.release()
    return out, None


def train_for_skew(ctx):
    """Traces of the soup, grid and job workloads -> a device-trained predictor."""
    from cpvulkan_amd import predictor as PR
    from cpvulkan_amd import workloads as W
    from cpvulkan_amd.workloads import DeviceMemory, job_desc, materialize
    mem = DeviceMemory(ctx)
    soup = materialize(W.soup_scene(1_000_000, 1920, 1080, seed=2), mem)
    grid = materialize(W.grid_scene(), mem)
    jobs = [job_desc(materialize(s, mem)) for s in W.job_scenes(64, seed=4)]
    traces = PR.collect_traces(ctx, passes=[soup, grid], job_sets=[jobs], configs=[(16, 4, 0), (16, 2, 0), (8, 1, 0)])
    mlp, report = PR.train_predictor(ctx, traces, samples=10_000, steps=3000)
    mem.release()
    report["traces"] = int(len(traces["cycles"]))
    return mlp, report


def timed(ctx, fn, steps: int, warmup: int) -> float:
    for _ in range(warmup):
        fn()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    ctx.synchronize()
    return steps / (time.perf_counter() - t0)


def bench_skewed(args, ctx, scene, workload, embedded: bool = False):
    """Plain single-draw path, counter-ordered persistent batch and the
    learned batch on the skewed frame.  embedded: the summary carried by
    the default line (no CPU baseline; the oracle check stays)."""
    import numpy as np
    from cpvulkan_amd import abi
    from cpvulkan_amd.workloads import DeviceMemory, job_desc, materialize
    mlp, report = train_for_skew(ctx)
    log(f"predictor trained: {report}")
    mem = DeviceMemory(ctx)
    m = materialize(scene, mem)
    cpu, check, refs = None, "skipped", None
    if not args.no_check or (not embedded and args.cpu_baseline != "none"):
        th = oracle_threads()
        mode = "sample" if (args.cpu_baseline == "sample" and not embedded) else "full"
        refs, dt = oracle_frames([scene], mode, threads=th)
        if not embedded and args.cpu_baseline != "none":
            cpu = cpu_entry(dt, mode, f"one frame ({scene.primitive_count} triangles)", threads=th)
    steps, warmup = args.steps, args.warmup
    # plain single-draw path: immediate render call, hardware-ordered tile grid
    ctx.set_config(tile_size=args.tile, waves_per_tile=args.waves, scheduler=abi.SCHED["GRID"], count_fragments=0)
    m.reset_attachments()
    ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
    ctx.synchronize()
    plain_img = (m.read_color(), m.read_ds(), None)
    plain_fps = timed(ctx, lambda: ctx.render_pass(m.pass_desc, m.pipelines, m.draws), steps, warmup)
    plain_raster = ctx.timing()[0]
    # the counter-ordered persistent queue, recorded: what the learned order is compared with
    ctx.set_config(tile_size=args.tile, waves_per_tile=args.waves, scheduler=abi.SCHED["COUNTER"], count_fragments=0)
    with ctx.batch([job_desc(m)]) as cbatch:
        counter_fps = timed(ctx, cbatch.submit, steps, warmup)
        cbatch.wait()
    # batched renderer: recorded hipGraphs, learned order, pipelined submits
    ctx.set_predictor(mlp)
    ctx.set_config(tile_size=args.tile, waves_per_tile=args.waves, scheduler=abi.SCHED["PREDICTED"], count_fragments=0)
    batch = ctx.batch([job_desc(m)])
    m.reset_attachments()
    batch.submit()
    batch.wait()
    sched_img = (m.read_color(), m.read_ds(), None)
    sched_fps = timed(ctx, batch.submit, steps, warmup)
    batch.wait()
    # the same batch one submit at a time (no pipelining): what the learned order alone gives
    one_ms = []
    for _ in range(max(steps // 2, 3)):
        batch.submit()
        batch.wait()
        one_ms.append(batch.timing())
    batch.close()
    ctx.set_predictor(None)
    mlp.close()
    if refs is not None and refs[0] is not None:
        ok = same(plain_img, refs[0]) and same(sched_img, refs[0])
        check = "bit-exact" if ok else "MISMATCH"
        if not ok:
            return None, dict(error="GPU frame differs from the CPU oracle")
    else:
        check = "learned-order frame identical to plain frame" if same(plain_img, sched_img) else "MISMATCH"
        if check == "MISMATCH":
            return None, dict(error="learned-order frame differs from the plain frame")
    b_fb = scene.color_bytes + scene.ds_bytes
    out = dict(
        value=round(sched_fps, 3), ms_per_step=round(1000.0 / sched_fps, 4),
        config=dict(workload=workload, width=scene.width, height=scene.height, triangles=scene.primitive_count,
                    tile=args.tile, waves_per_tile=args.waves, scheduler="predicted", batch="hipgraph"),
        plain_single_draw_fps=round(plain_fps, 3), counter_queue_batch_fps=round(counter_fps, 3),
        predicted_batch_fps=round(sched_fps, 3), speedup_vs_plain=round(sched_fps / plain_fps, 4),
        predicted_one_submit_ms=percentiles(one_ms), predictor=report,
        plain_raster_ms=round(plain_raster, 4),
        roofline=roofline(plain_raster, b_fb), cpu_baseline=cpu, check=check,
    )
    mem.release()
    return out, None


def bench_jobs(args, ctx):
    import numpy as np
    from cpvulkan_amd import workloads as W
    from cpvulkan_amd.workloads import DeviceMemory, job_desc, materialize
    scenes = W.job_scenes(256, seed=4)
    mem = DeviceMemory(ctx)
    ms = [materialize(s, mem) for s in scenes]
    jobs = [job_desc(m) for m in ms]
    tris = sum(s.primitive_count for s in scenes)
    table, best = ctx.autotune(jobs, repeats=2)
    log(f"autotune: {table} -> {best.tile_size}x{best.waves_per_tile}")
    best.count_fragments = 1
    ctx.apply_config(best)
    batch = ctx.batch(jobs)
    for m in ms:
        m.reset_attachments()
    batch.submit()
    fragments = batch.wait(count=True)
    batch.close()
    cpu, check = None, "skipped"
    if args.cpu_baseline != "none" or not args.no_check:
        mode = "sample" if args.cpu_baseline == "sample" else "full"
        refs, dt = oracle_frames(scenes, mode)
        cpu = cpu_entry(dt, mode, f"all {len(scenes)} jobs ({tris} triangles)")
        if refs[0] is not None and not args.no_check:
            ok = all(same((m.read_color(), m.read_ds(), None), r) for m, r in zip(ms, refs))
            ok = ok and fragments == sum(r[2] for r in refs)
            check = "bit-exact" if ok else "MISMATCH"
            if not ok:
                return None, dict(error="GPU frame differs from the CPU oracle")
    best.count_fragments = 0
    ctx.apply_config(best)
    batch = ctx.batch(jobs)
    for _ in range(args.warmup):
        batch.submit()
    batch.wait()
    gpu_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.submit()
        batch.wait()
        gpu_ms.append(batch.timing())
    wall = time.perf_counter() - t0
    fps = args.steps / wall
    batch.close()
    b_fb = sum(s.color_bytes + s.ds_bytes for s in scenes)
    out = dict(
        value=round(fps, 3), ms_per_step=round(wall * 1000.0 / args.steps, 4),
        config=dict(workload="jobs-256-zipf", jobs=len(scenes), triangles=tris, tile=best.tile_size,
                    waves_per_tile=best.waves_per_tile, batch="hipgraph", autotune=[list(t) for t in table]),
        fragments_per_frame=fragments, fragments_per_s=round(fragments * fps, 1),
        batch_gpu_ms=round(float(np.mean(gpu_ms)), 4),
        roofline=roofline(float(np.mean(gpu_ms)), b_fb), cpu_baseline=cpu, check=check,
    )
    mem.release()
    return out, None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="soup")
    ap.add_argument("--cpu-baseline", default="full", choices=["full", "sample", "none"])
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-skewed", action="store_true", help="leave the skewed-frame comparison out of the soup line")
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--waves", type=int, default=4)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the draw path runs on a single GPU (use --gpus 1)", file=sys.stderr)
        return 2
    import cpvulkan_amd
    ctx = cpvulkan_amd.Context(0)
    if args.copy_calibration:
        n = args.copy_calibration << 20
        src, dst = ctx.alloc(n), ctx.alloc(n)
        for _ in range(3):
            ctx.device_copy(dst, src, n)
        ctx.synchronize()
        print(json.dumps(dict(copy_bytes=n, launches=3)))
        ctx.close()
        return 0
    if args.workload == "jobs":
        out, err = bench_jobs(args, ctx)
    elif args.workload == "transfers":
        out, err = bench_transfers(args, ctx)
    else:
        scene, workload = make_scene(args.workload)
        log(f"workload {workload}: {scene.primitive_count} triangles, {scene.width}x{scene.height}")
        if args.workload == "skewed":
            out, err = bench_skewed(args, ctx, scene, workload)
        else:
            out, err = bench_single(args, ctx, scene, workload)
            if not err and args.workload == "soup" and not args.no_skewed:
                # the north-star comparison on the skewed 4K frame, measured in the same run
                sk_scene, sk_name = make_scene("skewed")
                log(f"skewed comparison: {sk_scene.primitive_count} triangles, {sk_scene.width}x{sk_scene.height}")
                sk, err = bench_skewed(args, ctx, sk_scene, sk_name, embedded=True)
                if not err:
                    out["skewed"] = {k: sk[k] for k in ("plain_single_draw_fps", "counter_queue_batch_fps",
                                                        "predicted_batch_fps", "speedup_vs_plain",
                                                        "predicted_one_submit_ms", "plain_raster_ms", "check")}
                    out["skewed"]["predictor_holdout_rank_corr"] = sk["predictor"].get("holdout_rank_corr")
    if err:
        print(json.dumps(dict(metric="frames/s", **err)), flush=True)
        return 1
    metric, unit = out.pop("metric", "frames/s"), out.pop("unit", "frames/s")
    line = dict(metric=metric, value=out.pop("value"), unit=unit, n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=out.pop("ms_per_step"), higher_is_better=True, scaling="weak",
                vs_baseline=None, dtype="f32", data="synthetic", config=out.pop("config"))
    line.update(out)
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
