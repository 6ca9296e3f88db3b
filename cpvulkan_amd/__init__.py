"""cpvulkan_amd -- MI355X-native draw path of the CPVulkan CPU Vulkan driver.

The product is native: libcpv.so (HIP kernels + the C-ABI of include/cpv.h)
and the Vulkan ICD built on it.  This package holds the build recipe, the
ctypes binding of the C-ABI (used by the benchmark and the tests) and the
synthetic workload generators of the benchmark configurations.

Nothing here falls back to a CPU path: loading the library or creating a
context on a machine without the built library or without a GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from . import abi

LIB_DIR = Path(__file__).resolve().parent / "lib"
LIBCPV = LIB_DIR / "libcpv.so"

_lib = None


def _declare(lib: C.CDLL) -> None:
    P = C.c_void_p
    lib.cpv_last_error.restype = C.c_char_p
    lib.cpv_context_create.argtypes = [C.c_int, C.POINTER(P)]
    lib.cpv_context_destroy.argtypes = [P]
    lib.cpv_context_destroy.restype = None
    lib.cpv_alloc.argtypes = [P, C.c_uint64, C.POINTER(C.c_uint64)]
    lib.cpv_free.argtypes = [P, C.c_uint64]
    lib.cpv_alloc_host.argtypes = [P, C.c_uint64, C.POINTER(C.c_uint64)]
    lib.cpv_free_host.argtypes = [P, C.c_uint64]
    lib.cpv_upload.argtypes = [P, C.c_uint64, P, C.c_uint64]
    lib.cpv_download.argtypes = [P, P, C.c_uint64, C.c_uint64]
    lib.cpv_synchronize.argtypes = [P]
    lib.cpv_device_copy.argtypes = [P, C.c_uint64, C.c_uint64, C.c_uint64]
    lib.cpv_copy_regions.argtypes = [P, C.POINTER(abi.CopyRegion), C.c_uint32]
    lib.cpv_fill_buffer.argtypes = [P, C.c_uint64, C.c_uint64, C.c_uint32]
    lib.cpv_update_buffer.argtypes = [P, C.c_uint64, P, C.c_uint64]
    lib.cpv_clear_rects.argtypes = [P, C.POINTER(abi.ImageLevel), C.POINTER(C.c_uint32), C.POINTER(abi.Rect), C.c_uint32]
    lib.cpv_blit_image.argtypes = [P, C.POINTER(abi.ImageLevel), C.POINTER(abi.ImageLevel), C.POINTER(abi.BlitRegion), C.c_uint32,
                                   C.c_uint32]
    lib.cpv_set_config.argtypes = [P, C.POINTER(abi.Config)]
    lib.cpv_get_config.argtypes = [P, C.POINTER(abi.Config)]
    lib.cpv_render_pass.argtypes = [P, C.POINTER(abi.PassDesc), C.POINTER(abi.PipelineDesc), C.POINTER(abi.DrawDesc),
                                    C.POINTER(C.c_uint32), C.c_uint32]
    lib.cpv_render_pass_noalloc.argtypes = lib.cpv_render_pass.argtypes
    lib.cpv_render_jobs.argtypes = [P, C.POINTER(abi.JobDesc), C.c_uint32]
    lib.cpv_last_fragment_count.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.cpv_last_timing.argtypes = [P, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.cpv_program_create_spirv.argtypes = [P, P, C.c_size_t, C.c_char_p, P, C.c_size_t, C.c_char_p, C.POINTER(P)]
    lib.cpv_program_destroy.argtypes = [P, P]
    lib.cpv_program_destroy.restype = None
    lib.cpv_program_source.argtypes = [P]
    lib.cpv_program_source.restype = C.c_char_p
    lib.cpv_abi_sizes.argtypes = [C.POINTER(C.c_uint32), C.c_uint32]
    lib.cpv_get_tile_traces.argtypes = [P, C.POINTER(abi.TileTrace), C.c_uint32, C.POINTER(C.c_uint32)]
    lib.cpv_batch_create.argtypes = [P, C.POINTER(abi.JobDesc), C.c_uint32, C.POINTER(P)]
    U32P = C.POINTER(C.c_uint32)
    lib.cpv_render_passes.argtypes = [P, C.POINTER(abi.PassDesc), C.c_uint32, C.POINTER(abi.PipelineDesc),
                                      C.POINTER(abi.DrawDesc), U32P, U32P]
    lib.cpv_batch_create_passes.argtypes = [P, C.POINTER(abi.PassDesc), C.c_uint32, C.POINTER(abi.PipelineDesc),
                                            C.POINTER(abi.DrawDesc), U32P, U32P, C.POINTER(P)]
    lib.cpv_launch_count.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.cpv_fence_create.argtypes = [P, C.c_int, C.POINTER(P)]
    lib.cpv_fence_signal.argtypes = [P]
    lib.cpv_fence_reset.argtypes = [P]
    lib.cpv_fence_status.argtypes = [P]
    lib.cpv_fence_wait.argtypes = [P, C.c_uint64]
    lib.cpv_fence_destroy.argtypes = [P]
    lib.cpv_fence_destroy.restype = None
    lib.cpv_batch_submit.argtypes = [P]
    lib.cpv_batch_wait.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.cpv_batch_timing.argtypes = [P, C.POINTER(C.c_float)]
    lib.cpv_batch_destroy.argtypes = [P]
    lib.cpv_batch_destroy.restype = None
    lib.cpv_autotune.argtypes = [P, C.POINTER(abi.JobDesc), C.c_uint32, C.c_uint32, C.POINTER(abi.AutotuneEntry), C.c_uint32,
                                 C.POINTER(C.c_uint32), C.POINTER(abi.Config)]
    FP = C.POINTER(C.c_float)
    lib.cpv_mlp_create.argtypes = [P, C.c_uint64, C.POINTER(P)]
    lib.cpv_mlp_destroy.argtypes = [P]
    lib.cpv_mlp_destroy.restype = None
    lib.cpv_mlp_get_params.argtypes = [P, FP, C.c_uint32]
    lib.cpv_mlp_set_params.argtypes = [P, FP, C.c_uint32]
    lib.cpv_mlp_predict.argtypes = [P, FP, C.c_uint32, FP]
    lib.cpv_mlp_gradients.argtypes = [P, FP, FP, C.c_uint32, FP, FP]
    lib.cpv_mlp_train.argtypes = [P, FP, FP, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, FP]
    lib.cpv_set_predictor.argtypes = [P, P]
    lib.cpv_spirv_translate.argtypes = [P, C.c_size_t, C.c_char_p, P, C.c_size_t, C.c_char_p, C.c_char_p, C.c_size_t,
                                        C.POINTER(C.c_size_t)]
    lib.cpv_spirv_translate_error.restype = C.c_char_p
    lib.cpv_spirv_build.argtypes = [P, C.c_size_t, C.c_char_p, P, C.c_size_t, C.c_char_p, C.POINTER(C.c_size_t)]


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libcpv.so (built by cpvulkan_amd.build) and check its ABI."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIBCPV
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run python -m cpvulkan_amd.build")
    lib = C.CDLL(str(p))
    _declare(lib)
    abi.check_layout(lib)
    if path is None:
        _lib = lib
    return lib


class CpvError(RuntimeError):
    pass


def _check(lib: C.CDLL, rc: int, what: str) -> None:
    if rc != 0:
        raise CpvError(f"{what} failed ({rc}): {lib.cpv_last_error().decode()}")


class Context:
    """One GPU context of libcpv (a HIP stream and the pipeline's work buffers)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        self.handle = C.c_void_p()
        _check(self.lib, self.lib.cpv_context_create(device, C.byref(self.handle)), "cpv_context_create")
        self._allocs: list[int] = []

    def close(self) -> None:
        if self.handle:
            for a in self._allocs:
                self.lib.cpv_free(self.handle, a)
            self._allocs.clear()
            self.lib.cpv_context_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # memory
    def alloc(self, nbytes: int) -> int:
        addr = C.c_uint64()
        _check(self.lib, self.lib.cpv_alloc(self.handle, max(int(nbytes), 1), C.byref(addr)), "cpv_alloc")
        self._allocs.append(addr.value)
        return addr.value

    def free(self, addr: int) -> None:
        self._allocs.remove(addr)
        _check(self.lib, self.lib.cpv_free(self.handle, addr), "cpv_free")

    def upload(self, addr: int, data) -> None:
        import numpy as np
        a = np.ascontiguousarray(data)
        _check(self.lib, self.lib.cpv_upload(self.handle, addr, a.ctypes.data, a.nbytes), "cpv_upload")

    def download(self, addr: int, nbytes: int):
        import numpy as np
        out = np.empty(int(nbytes), dtype=np.uint8)
        _check(self.lib, self.lib.cpv_download(self.handle, out.ctypes.data, addr, out.nbytes), "cpv_download")
        return out

    def synchronize(self) -> None:
        _check(self.lib, self.lib.cpv_synchronize(self.handle), "cpv_synchronize")

    def device_copy(self, dst: int, src: int, nbytes: int) -> None:
        _check(self.lib, self.lib.cpv_device_copy(self.handle, dst, src, nbytes), "cpv_device_copy")

    # transfers (enqueued on the context stream, in order with render passes)
    def copy_regions(self, regions) -> None:
        arr = (abi.CopyRegion * max(len(regions), 1))(*regions)
        _check(self.lib, self.lib.cpv_copy_regions(self.handle, arr, len(regions)), "cpv_copy_regions")

    def fill_buffer(self, dst: int, size: int, data: int) -> None:
        _check(self.lib, self.lib.cpv_fill_buffer(self.handle, dst, size, data & 0xFFFFFFFF), "cpv_fill_buffer")

    def update_buffer(self, dst: int, data) -> None:
        import numpy as np
        a = np.ascontiguousarray(data)
        _check(self.lib, self.lib.cpv_update_buffer(self.handle, dst, a.ctypes.data, a.nbytes), "cpv_update_buffer")

    def clear_rects(self, level: abi.ImageLevel, value_words, rects) -> None:
        v = (C.c_uint32 * 4)(*[int(w) & 0xFFFFFFFF for w in value_words])
        arr = (abi.Rect * max(len(rects), 1))(*rects)
        _check(self.lib, self.lib.cpv_clear_rects(self.handle, C.byref(level), v, arr, len(rects)), "cpv_clear_rects")

    def blit_image(self, src: abi.ImageLevel, dst: abi.ImageLevel, regions, filter_: int) -> None:
        arr = (abi.BlitRegion * max(len(regions), 1))(*regions)
        _check(self.lib, self.lib.cpv_blit_image(self.handle, C.byref(src), C.byref(dst), arr, len(regions), filter_),
               "cpv_blit_image")

    # configuration
    def set_config(self, tile_size=16, waves_per_tile=4, scheduler=0, split_threshold=0, record_costs=0,
                   count_fragments=1) -> None:
        cfg = abi.Config(tile_size, waves_per_tile, scheduler, split_threshold, record_costs, count_fragments)
        _check(self.lib, self.lib.cpv_set_config(self.handle, C.byref(cfg)), "cpv_set_config")

    # rendering
    def render_pass(self, pass_desc, pipelines, draws, draw_pipeline=None) -> None:
        n = len(draws)
        parr = (abi.PipelineDesc * max(len(pipelines), 1))(*pipelines)
        darr = (abi.DrawDesc * max(n, 1))(*draws)
        dp = (C.c_uint32 * max(n, 1))(*(draw_pipeline or [0] * n))
        rc = self.lib.cpv_render_pass(self.handle, C.byref(pass_desc), parr, darr, dp, n)
        _check(self.lib, rc, "cpv_render_pass")

    def render_pass_noalloc(self, pass_desc, pipelines, draws, draw_pipeline=None) -> None:
        """cpv_render_pass_noalloc: render_pass with every working buffer in
        the GPU's fixed scratch array (no allocation)."""
        n = len(draws)
        parr = (abi.PipelineDesc * max(len(pipelines), 1))(*pipelines)
        darr = (abi.DrawDesc * max(n, 1))(*draws)
        dp = (C.c_uint32 * max(n, 1))(*(draw_pipeline or [0] * n))
        rc = self.lib.cpv_render_pass_noalloc(self.handle, C.byref(pass_desc), parr, darr, dp, n)
        _check(self.lib, rc, "cpv_render_pass_noalloc")

    def render_jobs(self, jobs) -> None:
        arr = (abi.JobDesc * max(len(jobs), 1))(*jobs)
        _check(self.lib, self.lib.cpv_render_jobs(self.handle, arr, len(jobs)), "cpv_render_jobs")

    def fragment_count(self) -> int:
        v = C.c_uint64()
        _check(self.lib, self.lib.cpv_last_fragment_count(self.handle, C.byref(v)), "cpv_last_fragment_count")
        return v.value

    def timing(self) -> tuple[float, float]:
        r, t = C.c_float(), C.c_float()
        _check(self.lib, self.lib.cpv_last_timing(self.handle, C.byref(r), C.byref(t)), "cpv_last_timing")
        return r.value, t.value

    def config(self) -> abi.Config:
        cfg = abi.Config()
        _check(self.lib, self.lib.cpv_get_config(self.handle, C.byref(cfg)), "cpv_get_config")
        return cfg

    def apply_config(self, cfg: abi.Config) -> None:
        _check(self.lib, self.lib.cpv_set_config(self.handle, C.byref(cfg)), "cpv_set_config")

    def tile_traces(self):
        """Per-tile cost traces of the last launch (record_costs on): a dict
        of numpy arrays features [n, 12], cycles [n], job [n], tile [n]."""
        import numpy as np
        n = C.c_uint32()
        _check(self.lib, self.lib.cpv_get_tile_traces(self.handle, None, 0, C.byref(n)), "cpv_get_tile_traces")
        arr = (abi.TileTrace * max(n.value, 1))()
        _check(self.lib, self.lib.cpv_get_tile_traces(self.handle, arr, n.value, C.byref(n)), "cpv_get_tile_traces")
        raw = np.frombuffer(arr, dtype=np.uint8).reshape(max(n.value, 1), C.sizeof(abi.TileTrace))[: n.value]
        f = raw[:, :52].copy().view(np.float32)
        u = raw[:, 52:64].copy().view(np.uint32)
        return dict(features=f[:, :12].copy(), cycles=f[:, 12].copy(), job=u[:, 0].copy(), tile=u[:, 1].copy())

    def batch(self, jobs) -> "Batch":
        return Batch(self, jobs)

    def batch_passes(self, passes) -> "Batch":
        """A recorded batch of render passes: passes is a list of
        (pass_desc, pipelines, draws, draw_pipeline) as for render_pass."""
        return Batch(self, None, passes=passes)

    def render_passes(self, passes) -> None:
        """cpv_render_passes over a list of (pass_desc, pipelines, draws,
        draw_pipeline) tuples, each as render_pass takes them."""
        args = _pass_arrays(passes)
        _check(self.lib, self.lib.cpv_render_passes(self.handle, *args), "cpv_render_passes")

    def launch_count(self) -> int:
        v = C.c_uint64()
        _check(self.lib, self.lib.cpv_launch_count(self.handle, C.byref(v)), "cpv_launch_count")
        return v.value

    def fence(self, signaled: bool = False) -> "Fence":
        return Fence(self, signaled)

    def autotune(self, jobs, repeats: int = 3):
        """Sweeps the tile configurations over a recorded batch of jobs;
        returns ([(tile, waves, ms)], best abi.Config)."""
        arr = (abi.JobDesc * max(len(jobs), 1))(*jobs)
        entries = (abi.AutotuneEntry * 16)()
        n = C.c_uint32()
        best = abi.Config()
        rc = self.lib.cpv_autotune(self.handle, arr, len(jobs), repeats, entries, 16, C.byref(n), C.byref(best))
        _check(self.lib, rc, "cpv_autotune")
        return [(e.tile_size, e.waves_per_tile, e.ms) for e in entries[: n.value]], best

    def set_predictor(self, mlp: "Mlp | None") -> None:
        _check(self.lib, self.lib.cpv_set_predictor(self.handle, mlp.handle if mlp else None), "cpv_set_predictor")


def _words(words):
    import numpy as np
    return np.ascontiguousarray(np.asarray(words, dtype=np.uint32))


def spirv_translate(vs_words, fs_words, vs_entry: str = "main", fs_entry: str = "main") -> str:
    """HIP source of a SPIR-V (vertex, fragment) pair (no device needed)."""
    lib = load_library()
    v, f = _words(vs_words), _words(fs_words)
    n = C.c_size_t()
    rc = lib.cpv_spirv_translate(v.ctypes.data, v.size, vs_entry.encode(), f.ctypes.data, f.size, fs_entry.encode(), None,
                                 0, C.byref(n))
    if rc != 0:
        raise CpvError(f"cpv_spirv_translate failed ({rc}): {lib.cpv_spirv_translate_error().decode()}")
    buf = C.create_string_buffer(n.value + 1)
    lib.cpv_spirv_translate(v.ctypes.data, v.size, vs_entry.encode(), f.ctypes.data, f.size, fs_entry.encode(), buf,
                            n.value + 1, C.byref(n))
    return buf.value.decode()


def spirv_build(vs_words, fs_words, vs_entry: str = "main", fs_entry: str = "main") -> int:
    """Translate and compile with hiprtc (no device needed); returns the code object size."""
    lib = load_library()
    v, f = _words(vs_words), _words(fs_words)
    n = C.c_size_t()
    rc = lib.cpv_spirv_build(v.ctypes.data, v.size, vs_entry.encode(), f.ctypes.data, f.size, fs_entry.encode(), C.byref(n))
    if rc != 0:
        raise CpvError(f"cpv_spirv_build failed ({rc}): {lib.cpv_spirv_translate_error().decode()}")
    return n.value


class Program:
    """A shader program translated from SPIR-V and compiled with hiprtc
    (cpv_program_create_spirv); set it as a pipeline's `program`."""

    def __init__(self, ctx: "Context", vs_words, fs_words, vs_entry: str = "main", fs_entry: str = "main"):
        self.ctx = ctx
        self.lib = ctx.lib
        v, f = _words(vs_words), _words(fs_words)
        self.handle = C.c_void_p()
        rc = self.lib.cpv_program_create_spirv(ctx.handle, v.ctypes.data, v.size, vs_entry.encode(), f.ctypes.data, f.size,
                                               fs_entry.encode(), C.byref(self.handle))
        _check(self.lib, rc, "cpv_program_create_spirv")

    @property
    def source(self) -> str:
        return self.lib.cpv_program_source(self.handle).decode()

    def close(self) -> None:
        if self.handle:
            self.lib.cpv_program_destroy(self.ctx.handle, self.handle)
            self.handle = C.c_void_p()


def _pass_arrays(passes):
    """The flat arrays of cpv_render_passes: every pass's pipelines and draws
    concatenated, draw_pipeline re-indexed, first_draw offsets."""
    pds, pipes, draws, dp, first = [], [], [], [], [0]
    for pass_desc, pipelines, pass_draws, draw_pipeline in passes:
        base = len(pipes)
        pds.append(pass_desc)
        pipes.extend(pipelines)
        draws.extend(pass_draws)
        dp.extend(base + k for k in (draw_pipeline or [0] * len(pass_draws)))
        first.append(len(draws))
    return ((abi.PassDesc * max(len(pds), 1))(*pds), len(pds), (abi.PipelineDesc * max(len(pipes), 1))(*pipes),
            (abi.DrawDesc * max(len(draws), 1))(*draws), (C.c_uint32 * max(len(dp), 1))(*dp),
            (C.c_uint32 * len(first))(*first))


class Fence:
    """cpv_fence_*: the end of the work enqueued on a context up to signal()."""

    def __init__(self, ctx: Context, signaled: bool = False):
        self.lib = ctx.lib
        self.handle = C.c_void_p()
        _check(self.lib, self.lib.cpv_fence_create(ctx.handle, 1 if signaled else 0, C.byref(self.handle)), "cpv_fence_create")

    def signal(self) -> None:
        _check(self.lib, self.lib.cpv_fence_signal(self.handle), "cpv_fence_signal")

    def reset(self) -> None:
        _check(self.lib, self.lib.cpv_fence_reset(self.handle), "cpv_fence_reset")

    def status(self) -> int:
        """abi.OK, or abi.NOT_READY."""
        rc = self.lib.cpv_fence_status(self.handle)
        if rc < 0:
            _check(self.lib, rc, "cpv_fence_status")
        return rc

    def wait(self, timeout_ns: int = (1 << 64) - 1) -> int:
        """abi.OK, or abi.TIMEOUT."""
        rc = self.lib.cpv_fence_wait(self.handle, timeout_ns)
        if rc < 0:
            _check(self.lib, rc, "cpv_fence_wait")
        return rc

    def close(self) -> None:
        if self.handle:
            self.lib.cpv_fence_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """A recorded set of render jobs, or of render passes, replayed as
    hipGraphs (cpv_batch_*)."""

    def __init__(self, ctx: Context, jobs, passes=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.handle = C.c_void_p()
        if passes is not None:
            self._arrays = _pass_arrays(passes)
            _check(self.lib, self.lib.cpv_batch_create_passes(ctx.handle, *self._arrays, C.byref(self.handle)),
                   "cpv_batch_create_passes")
            return
        self._jobs = (abi.JobDesc * max(len(jobs), 1))(*jobs)
        _check(self.lib, self.lib.cpv_batch_create(ctx.handle, self._jobs, len(jobs), C.byref(self.handle)), "cpv_batch_create")

    def submit(self) -> None:
        _check(self.lib, self.lib.cpv_batch_submit(self.handle), "cpv_batch_submit")

    def wait(self, count: bool = False):
        v = C.c_uint64()
        _check(self.lib, self.lib.cpv_batch_wait(self.handle, C.byref(v) if count else None), "cpv_batch_wait")
        return v.value if count else None

    def timing(self) -> float:
        ms = C.c_float()
        _check(self.lib, self.lib.cpv_batch_timing(self.handle, C.byref(ms)), "cpv_batch_timing")
        return ms.value

    def close(self) -> None:
        if self.handle:
            self.lib.cpv_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Mlp:
    """The tile-cost predictor (cpv_mlp_*): BF16 MLP 12 -> 64 -> 64 -> 1 run
    and trained on the GPU."""

    def __init__(self, ctx: Context, seed: int = 0):
        self.ctx = ctx
        self.lib = ctx.lib
        self.handle = C.c_void_p()
        _check(self.lib, self.lib.cpv_mlp_create(ctx.handle, seed, C.byref(self.handle)), "cpv_mlp_create")

    @staticmethod
    def _f(a):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=np.float32)
        return a, a.ctypes.data_as(C.POINTER(C.c_float))

    def params(self):
        import numpy as np
        out = np.empty(abi.MLP_PARAMS, np.float32)
        _check(self.lib, self.lib.cpv_mlp_get_params(self.handle, self._f(out)[1], abi.MLP_PARAMS), "cpv_mlp_get_params")
        return out

    def set_params(self, p) -> None:
        a, ptr = self._f(p)
        _check(self.lib, self.lib.cpv_mlp_set_params(self.handle, ptr, a.size), "cpv_mlp_set_params")

    def predict(self, x):
        import numpy as np
        xa, xp = self._f(x)
        rows = xa.shape[0]
        y = np.empty(rows, np.float32)
        _check(self.lib, self.lib.cpv_mlp_predict(self.handle, xp, rows, self._f(y)[1]), "cpv_mlp_predict")
        return y

    def gradients(self, x, t):
        import numpy as np
        xa, xp = self._f(x)
        ta, tp = self._f(t)
        g = np.empty(abi.MLP_PARAMS, np.float32)
        loss = C.c_float()
        rc = self.lib.cpv_mlp_gradients(self.handle, xp, tp, xa.shape[0], self._f(g)[1], C.byref(loss))
        _check(self.lib, rc, "cpv_mlp_gradients")
        return g, loss.value

    def train(self, x, t, steps: int, batch: int = 1024, lr: float = 1e-3, seed: int = 0) -> float:
        xa, xp = self._f(x)
        ta, tp = self._f(t)
        loss = C.c_float()
        rc = self.lib.cpv_mlp_train(self.handle, xp, tp, xa.shape[0], steps, batch, lr, seed, C.byref(loss))
        _check(self.lib, rc, "cpv_mlp_train")
        return loss.value

    def close(self) -> None:
        if self.handle:
            self.lib.cpv_mlp_destroy(self.handle)
            self.handle = C.c_void_p()
