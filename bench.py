"""Draw-path benchmark: frames/s of one workload on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload soup|cube|grid|skewed]
                    [--cpu-baseline full|sample|none] [--no-check]

A step is one frame: the render pass of the workload (clear, vertex stage,
setup, binning, raster/fragment) over inputs already resident in HBM.  The
default workload is the 1920x1080, 1M-triangle alpha-blended soup
(R8G8B8A8_UNORM + D32_SFLOAT, LESS, SRC_ALPHA/ONE_MINUS_SRC_ALPHA).

Before timing, the GPU frame is compared byte for byte with the CPU oracle's
frame; the oracle's single-threaded wall time for that frame is the
reported CPU baseline.  Prints one JSON line.
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="soup")
    ap.add_argument("--cpu-baseline", default="full", choices=["full", "sample", "none"])
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--waves", type=int, default=4)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the draw path runs on a single GPU (use --gpus 1)", file=sys.stderr)
        return 2
    if args.copy_calibration:
        import cpvulkan_amd
        ctx = cpvulkan_amd.Context(0)
        n = args.copy_calibration << 20
        src, dst = ctx.alloc(n), ctx.alloc(n)
        for _ in range(3):
            ctx.device_copy(dst, src, n)
        ctx.synchronize()
        print(json.dumps(dict(copy_bytes=n, launches=3)))
        ctx.close()
        return 0

    import numpy as np

    import cpvulkan_amd
    from cpvulkan_amd import abi
    from cpvulkan_amd.workloads import DeviceMemory, HostMemory, materialize

    scene, workload = make_scene(args.workload)
    log(f"workload {workload}: {scene.primitive_count} triangles, {scene.width}x{scene.height}")

    ctx = cpvulkan_amd.Context(0)
    ctx.set_config(tile_size=args.tile, waves_per_tile=args.waves)
    mem = DeviceMemory(ctx)
    m = materialize(scene, mem)

    # one frame with exact fragment counting (the count is a statistic; it
    # forces coverage evaluation of every pixel), then one frame the way the
    # timed frames run (counting off, hierarchical depth culling on); both
    # go through the correctness gate
    ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
    ctx.synchronize()
    fragments = ctx.fragment_count()
    gpu_color, gpu_ds = m.read_color(), m.read_ds()
    ctx.set_config(tile_size=args.tile, waves_per_tile=args.waves, count_fragments=0)
    m.reset_attachments()
    ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
    ctx.synchronize()
    fast_color, fast_ds = m.read_color(), m.read_ds()
    log(f"gpu frames done: {fragments} fragments")

    cpu = None
    check = "skipped"
    if args.cpu_baseline != "none" or not args.no_check:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib
        hm = materialize(scene, HostMemory())
        prims = scene.primitive_count
        if args.cpu_baseline == "sample":
            n = min(prims, 50_000)
            t0 = time.perf_counter()
            oracle_lib.render(hm, (0, n))
            dt = time.perf_counter() - t0
            frame_s = dt * prims / n
            cpu = dict(value=1.0 / frame_s, unit="frames/s", cores=1, kind="port",
                       sample=f"first {n} of {prims} triangles, frame time scaled by triangle count; "
                              f"{cpu_model()}")
        else:
            t0 = time.perf_counter()
            ref_frags = oracle_lib.render(hm)
            dt = time.perf_counter() - t0
            cpu = dict(value=1.0 / dt, unit="frames/s", cores=1, kind="port",
                       sample=f"one full frame ({prims} triangles) single-threaded, {dt:.2f} s; {cpu_model()}")
            if not args.no_check:
                ref_color, ref_ds = hm.read_color(), hm.read_ds()
                same = (np.array_equal(ref_color, gpu_color) and np.array_equal(ref_ds, gpu_ds)
                        and np.array_equal(ref_color, fast_color) and np.array_equal(ref_ds, fast_ds)
                        and ref_frags == fragments)
                check = "bit-exact" if same else "MISMATCH"
                log(f"oracle frame {dt:.2f} s, check {check}")
                if not same:
                    print(json.dumps(dict(metric="frames/s", error="GPU frame differs from the CPU oracle")))
                    return 1

    # timed frames
    for _ in range(args.warmup):
        ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
    ctx.synchronize()
    raster_ms = []
    total_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.render_pass(m.pass_desc, m.pipelines, m.draws)
        r, t = ctx.timing()
        raster_ms.append(r)
        total_ms.append(t)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1000.0 / args.steps
    fps = args.steps / wall

    b_fb = scene.color_bytes + scene.ds_bytes
    raster_avg = float(np.mean(raster_ms))
    achieved = b_fb / (raster_avg * 1e-3) / 1e9
    out = dict(
        metric="frames/s", value=round(fps, 3), unit="frames/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=round(ms_per_step, 4), higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f32", data="synthetic",
        config=dict(workload=workload, width=scene.width, height=scene.height, triangles=scene.primitive_count,
                    tile=args.tile, waves_per_tile=args.waves, color=fmt_name(scene.color_format),
                    depth=fmt_name(scene.ds_format)),
        fragments_per_frame=fragments, fragments_per_s=round(fragments * fps, 1),
        raster_ms=round(raster_avg, 4), frame_gpu_ms=round(float(np.mean(total_ms)), 4),
        roofline=dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(achieved / HBM_PEAK_GBS, 5), traffic=None),
        cpu_baseline=cpu, check=check,
    )
    print(json.dumps(out), flush=True)
    mem.release()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
