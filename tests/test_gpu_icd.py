"""The draw path behind the Vulkan ICD entry points (GPU only): a Vulkan
application (tests/icd/icd_cube.cpp) renders the 15-draw_cube frame through
vk_icdGetInstanceProcAddr -- SPIR-V shader modules, a graphics pipeline, a
uniform-buffer descriptor set, vkCmdDraw recorded into a command buffer and
replayed by vkQueueSubmit -- and its colour and depth images must equal the
CPU oracle's bytes."""
import copy
import os
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib
from cpvulkan_amd import abi
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


def _run_cube(tmp_path, size, mode):
    np.asarray(S.vs_mvp_color(), np.uint32).tofile(tmp_path / "vs.spv")
    np.asarray(S.fs_color(), np.uint32).tofile(tmp_path / "fs.spv")
    W.cube_vertices().tofile(tmp_path / "vertices.bin")
    W.cube_mvp(size, size).tofile(tmp_path / "mvp.bin")
    print(f"[icd] running {APP.name} {size}x{size} {mode}", flush=True)
    res = subprocess.run([str(APP), str(ICD), str(tmp_path), str(size), str(size), mode], capture_output=True, text=True,
                         timeout=240)
    print(res.stdout + res.stderr, flush=True)
    assert res.returncode == 0
    return np.fromfile(tmp_path / "color.bin", np.uint8), np.fromfile(tmp_path / "depth.bin", np.uint8)


def test_draw_indirect_through_the_icd(tmp_path):
    """vkCmdDrawIndirect with two records covering the cube's 36 vertices
    draws the same frame as vkCmdDraw(36)."""
    color, depth = _run_cube(tmp_path, 200, "indirect")
    ref_color, ref_ds, _ = oracle_lib.render_scene(W.cube_scene(width=200, height=200))
    assert np.array_equal(color, ref_color) and np.array_equal(depth, ref_ds)


def test_lines_through_the_icd(tmp_path):
    """The cube's vertices as a LINE_LIST with lineWidth 3 from the
    pipeline's rasterization state (ProcessLines through the ICD)."""
    color, depth = _run_cube(tmp_path, 180, "lines")
    scene = W.cube_scene(width=180, height=180)
    scene.pipeline["topology"] = abi.TOPOLOGY["LINE_LIST"]
    scene.pipeline["line_width"] = 3.0
    ref_color, ref_ds, _ = oracle_lib.render_scene(scene)
    px = ref_color.reshape(-1, 4)
    assert (px != px[0]).any(axis=1).sum() > 100      # the lines cover pixels
    assert np.array_equal(color, ref_color), f"colour differs in {np.count_nonzero(color != ref_color)} bytes"
    assert np.array_equal(depth, ref_ds), f"depth differs in {np.count_nonzero(depth != ref_ds)} bytes"


def test_dynamic_blend_constants_and_depth_bounds(tmp_path):
    """vkCmdSetBlendConstants / vkCmdSetDepthBounds between two draws of a
    pipeline that declares both dynamic: the oracle draws the same with two
    pipelines holding those values statically."""
    size = 160
    color, depth = _run_cube(tmp_path, size, "dynamic")
    scene = W.cube_scene(width=size, height=size)
    blend = dict(src_color=abi.BLEND["CONSTANT_COLOR"], dst_color=abi.BLEND["ONE_MINUS_CONSTANT_COLOR"], color_op=0,
                 src_alpha=abi.BLEND["CONSTANT_ALPHA"], dst_alpha=abi.BLEND["ONE_MINUS_CONSTANT_ALPHA"], alpha_op=0)
    pipes = []
    for k, bounds in (((0.25, 0.5, 0.75, 1.0), (0.0, 1.0)), ((0.8, 0.1, 0.3, 0.5), (0.0, 0.995))):
        p = dict(scene.pipeline, blend=blend, blend_constants=k, depth_bounds_test=1, depth_bounds=bounds)
        pipes.append(p)
    scene.pipelines = pipes
    scene.draws = [scene.draws[0], copy.deepcopy(scene.draws[0])]
    scene.draw_pipeline = [0, 1]
    ref_color, ref_ds, _ = oracle_lib.render_scene(scene)
    assert np.array_equal(color, ref_color), f"colour differs in {np.count_nonzero(color != ref_color)} bytes"
    assert np.array_equal(depth, ref_ds), f"depth differs in {np.count_nonzero(depth != ref_ds)} bytes"


def test_clear_attachments_inside_a_render_pass(tmp_path):
    """Draw, vkCmdClearAttachments (colour and depth rectangles), draw again:
    the oracle renders the cube, clears the rectangles, and renders the cube
    again over the result (LOAD)."""
    size = 240
    color, depth = _run_cube(tmp_path, size, "midclear")
    scene = W.cube_scene(width=size, height=size)
    c1, d1, _ = oracle_lib.render_scene(scene)
    rect = [abi.Rect(size // 4, size // 4, size // 2, size // 3)]
    oracle_lib.clear_rects(c1, abi.FORMAT["B8G8R8A8_UNORM"], size, size, [0x3F800000, 0, 0x3F000000, 0x3F800000], rect)
    oracle_lib.clear_rects(d1, abi.FORMAT["D16_UNORM"], size, size, [0x3F000000, 0, 0, 0], rect)
    again = copy.deepcopy(scene)
    again.color_load = again.ds_load = abi.LOAD_OP["LOAD"]
    again.initial_color, again.initial_ds = c1, d1
    c2, d2, _ = oracle_lib.render_scene(again)
    assert np.array_equal(color, c2), f"colour differs in {np.count_nonzero(color != c2)} bytes"
    assert np.array_equal(depth, d2), f"depth differs in {np.count_nonzero(depth != d2)} bytes"


@pytest.mark.parametrize("mode", ["indexed_indirect", "indirect_count", "indexed_indirect_count"])
def test_indexed_and_counted_indirect_through_the_icd(tmp_path, mode):
    """vkCmdDrawIndexedIndirect, vkCmdDrawIndirectCountKHR and
    vkCmdDrawIndexedIndirectCountAMD split the cube's 36 vertices into two
    records; the counted forms read a count of 2 with maxDrawCount 3, so
    their third record (which would draw more) is ignored: the frame is
    vkCmdDraw(36)'s."""
    color, depth = _run_cube(tmp_path, 200, mode)
    ref_color, ref_ds, _ = oracle_lib.render_scene(W.cube_scene(width=200, height=200))
    assert np.array_equal(color, ref_color), f"colour differs in {np.count_nonzero(color != ref_color)} bytes"
    assert np.array_equal(depth, ref_ds), f"depth differs in {np.count_nonzero(depth != ref_ds)} bytes"


def test_dynamic_stencil_masks_and_references(tmp_path):
    """vkCmdSetStencilCompareMask / WriteMask / Reference between two draws of
    a pipeline with EQUAL / REPLACE stencil on a D24_UNORM_S8_UINT
    attachment (separate front and back references in the second draw): the
    oracle draws the same with two pipelines holding those values
    statically (CommandBuffer.h:40-42)."""
    size = 180
    color, depth = _run_cube(tmp_path, size, "stencil")
    scene = W.cube_scene(width=size, height=size)
    scene.ds_format = abi.FORMAT["D24_UNORM_S8_UINT"]
    S_ = abi.STENCIL
    pipes = []
    for (cmask, wmask, ref_front, ref_back) in ((0x00, 0xF0, 0xA5, 0xA5), (0xF0, 0x0F, 0xAF, 0xA3)):
        p = dict(scene.pipeline, stencil_test=1,
                 front=(S_["KEEP"], S_["REPLACE"], S_["KEEP"], abi.COMPARE["EQUAL"], cmask, wmask, ref_front),
                 back=(S_["KEEP"], S_["REPLACE"], S_["KEEP"], abi.COMPARE["EQUAL"], cmask, wmask, ref_back))
        pipes.append(p)
    scene.pipelines = pipes
    scene.draws = [scene.draws[0], copy.deepcopy(scene.draws[0])]
    scene.draw_pipeline = [0, 1]
    ref_color, ref_ds, _ = oracle_lib.render_scene(scene)
    st = ref_ds.reshape(-1, 4)[:, 3]
    assert 0xAF in st and 0 in st, "the second draw's front faces replace the stencil inside the cube"
    assert np.array_equal(color, ref_color), f"colour differs in {np.count_nonzero(color != ref_color)} bytes"
    assert np.array_equal(depth, ref_ds), f"depth/stencil differs in {np.count_nonzero(depth != ref_ds)} bytes"


def _mvps(size):
    """The sample's MVP, then the same turned about z by 0.3 and 0.7 radians."""
    base = W.cube_mvp(size, size).reshape(4, 4)
    out = [base]
    for a in (0.3, 0.7):
        c, s = np.float32(np.cos(a)), np.float32(np.sin(a))
        rot = np.eye(4, dtype=np.float32)
        rot[0, 0], rot[0, 1], rot[1, 0], rot[1, 1] = c, -s, s, c
        out.append((base @ rot).astype(np.float32))
    return out


def _cube_pass(size, clear, mvp):
    s = W.cube_scene(width=size, height=size)
    s.clear_color = (clear,) * 4
    s.draws[0].uniforms = {0: np.ascontiguousarray(mvp, np.float32).reshape(-1).view(np.uint8)}
    return s


def _run_traced(tmp_path, size, mode, mvps):
    np.asarray(S.vs_mvp_color(), np.uint32).tofile(tmp_path / "vs.spv")
    np.asarray(S.fs_color(), np.uint32).tofile(tmp_path / "fs.spv")
    W.cube_vertices().tofile(tmp_path / "vertices.bin")
    W.cube_mvp(size, size).tofile(tmp_path / "mvp.bin")
    for j, m in enumerate(mvps):
        np.ascontiguousarray(m, np.float32).tofile(tmp_path / f"mvp{j}.bin")
    env = dict(os.environ, CPVULKAN_AMD_TRACE_SUBMITS="1")
    print(f"[icd] running {APP.name} {size}x{size} {mode}", flush=True)
    res = subprocess.run([str(APP), str(ICD), str(tmp_path), str(size), str(size), mode], capture_output=True, text=True,
                         timeout=240, env=env)
    print(res.stdout + res.stderr, flush=True)
    assert res.returncode == 0
    return res


def _check_passes(tmp_path, size, sub, mvp):
    for k in range(3):
        clear = float(np.float32(0.2) + np.float32(0.1) * np.float32(k))
        ref_color, ref_ds, _ = oracle_lib.render_scene(_cube_pass(size, clear, mvp))
        color = np.fromfile(tmp_path / f"color_{sub}_{k}.bin", np.uint8)
        depth = np.fromfile(tmp_path / f"depth_{sub}_{k}.bin", np.uint8)
        assert np.array_equal(color, ref_color), f"submit {sub} pass {k}: colour differs in {np.count_nonzero(color != ref_color)} bytes"
        assert np.array_equal(depth, ref_ds), f"submit {sub} pass {k}: depth differs"


def test_three_render_passes_one_launch(tmp_path):
    """Three render passes of one command buffer, each into its own
    framebuffer with its own clear colour, go to the GPU as one launch."""
    size = 160
    res = _run_traced(tmp_path, size, "multipass", [])
    assert re.findall(r"submit (\d+): (\d+) launches", res.stderr) == [("1", "1")]
    _check_passes(tmp_path, size, 0, W.cube_mvp(size, size).reshape(4, 4))


def test_resubmitted_command_buffer_replays_its_batch(tmp_path):
    """A command buffer recorded without ONE_TIME_SUBMIT, submitted three
    times with the uniform rewritten in between: every submit is one launch
    of the recorded batch and draws the frame of its own MVP; the fence
    reports VK_NOT_READY before any submit and after a reset, VK_TIMEOUT for
    a timed wait that nothing will signal, and waits for any or all of two
    fences (Fence.cpp:47-56, 157-200)."""
    size = 144
    mvps = _mvps(size)
    res = _run_traced(tmp_path, size, "resubmit", mvps)
    assert re.findall(r"submit (\d+): (\d+) launches", res.stderr) == [("1", "1"), ("2", "1"), ("3", "1")]
    fence = dict(re.findall(r"fence ([a-z0-9 ]+): (-?\d+)", res.stdout))
    assert fence["unsubmitted status"] == "1"        # VK_NOT_READY
    assert fence["unsubmitted wait 1ms"] == "2"      # VK_TIMEOUT
    assert fence["any of two"] == "0" and fence["all of two"] == "2"
    assert fence["after wait"] == "0" and fence["after reset"] == "1"
    for j, m in enumerate(mvps):
        _check_passes(tmp_path, size, j, m)
