"""The draw path behind the Vulkan ICD entry points (GPU only): a Vulkan
application (tests/icd/icd_cube.cpp) renders the 15-draw_cube frame through
vk_icdGetInstanceProcAddr -- SPIR-V shader modules, a graphics pipeline, a
uniform-buffer descriptor set, vkCmdDraw recorded into a command buffer and
replayed by vkQueueSubmit -- and its colour and depth images must equal the
CPU oracle's bytes."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib
from cpvulkan_amd import spirv as S
from cpvulkan_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
APP = ROOT / "tests" / "icd" / "build" / "icd_cube"
ICD = ROOT / "cpvulkan_amd" / "lib" / "libcpvulkan_amd_icd.so"


@pytest.mark.parametrize("size", [500, 333])
def test_cube_through_the_icd(tmp_path, size):
    np.asarray(S.vs_mvp_color(), np.uint32).tofile(tmp_path / "vs.spv")
    np.asarray(S.fs_color(), np.uint32).tofile(tmp_path / "fs.spv")
    W.cube_vertices().tofile(tmp_path / "vertices.bin")
    W.cube_mvp(size, size).tofile(tmp_path / "mvp.bin")
    print(f"[icd] running {APP.name} {size}x{size}", flush=True)
    res = subprocess.run([str(APP), str(ICD), str(tmp_path), str(size), str(size)], capture_output=True, text=True, timeout=240)
    print(res.stdout + res.stderr, flush=True)
    assert res.returncode == 0
    color = np.fromfile(tmp_path / "color.bin", np.uint8)
    depth = np.fromfile(tmp_path / "depth.bin", np.uint8)
    ref_color, ref_ds, _ = oracle_lib.render_scene(W.cube_scene(width=size, height=size))
    assert np.array_equal(color, ref_color), f"colour differs in {np.count_nonzero(color != ref_color)} bytes"
    assert np.array_equal(depth, ref_ds), f"depth differs in {np.count_nonzero(depth != ref_ds)} bytes"
