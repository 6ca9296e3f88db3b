"""The tile-parallel oracle (cpvref_render_pass_parallel, OpenMP over 64x64
pixel tiles) against the single-threaded replay it restates: colour bytes,
depth/stencil bytes and covered-fragment counts identical for every
topology and fixed-function state the GPU tests cover, including tiles cut
through the middle of primitives, viewports smaller than the attachment and
odd extents.  The parallel leg is the CPU baseline's multi-core number and
the checker of the full-size GPU parity tests, so it must equal the
reference-faithful serial leg bit for bit."""
import numpy as np
import pytest

import oracle_lib
import scenes as S
from cpvulkan_amd import abi
from cpvulkan_amd import workloads as W

F, CMP, ST = abi.FORMAT, abi.COMPARE, abi.STENCIL


def _same(scene, threads=7):
    a = oracle_lib.render_scene(scene)
    b = oracle_lib.render_scene(scene, threads=threads)
    assert np.array_equal(a[0], b[0]), f"{scene.name}: colour differs"
    assert np.array_equal(a[1], b[1]), f"{scene.name}: depth/stencil differs"
    assert a[2] == b[2], f"{scene.name}: fragments {a[2]} != {b[2]}"
    return a[2]


CASES = {
    "soup": lambda: W.soup_scene(60_000, 333, 197, seed=3),
    "skewed": lambda: W.skewed_scene(80_000, 640, 360, seed=5),
    "cube": lambda: W.cube_scene(),
    "cube_back_culled": lambda: W.cube_scene(cull="BACK"),
    "grid_textured": lambda: W.grid_scene(48, 301, 157, tex_size=64),
    "strip": lambda: S.strip(400, 200, 130, seed=2),
    "fan": lambda: S.strip(300, 200, 130, seed=2, fan=True),
    "indexed_u16": lambda: S.indexed(S.soup(2000, 190, 150, seed=4), index_dtype=np.uint16, vertex_offset=3),
    "instanced": lambda: S.instanced(300, 190, 150, seed=1),
    "perspective_w": lambda: S.soup(3000, 170, 130, seed=8, w_range=(0.5, 3.0)),
    "big_offscreen": lambda: S.soup(400, 211, 149, seed=9, size_px=np.log(90.0), offscreen=0.4),
    "viewport": lambda: S.with_viewport(S.soup(3000, 200, 150, seed=5), (0.0, 0.0, 137.0, 91.0, 0.25, 0.75)),
    "load_op": lambda: S.load_op(S.soup(3000, 160, 120, seed=6), seed=2),
    "multi_draw": lambda: S.multi_draw(seed=3),
    "stencil": lambda: S.with_formats(
        S.with_pipeline(S.soup(3000, 160, 120, seed=6), stencil_test=1,
                        front=(ST["KEEP"], ST["INC_CLAMP"], ST["DEC_CLAMP"], CMP["ALWAYS"], 0xFF, 0xFF, 1),
                        back=(ST["ZERO"], ST["REPLACE"], ST["INVERT"], CMP["GREATER"], 0x0F, 0xF0, 200),
                        depth_compare=CMP["LESS_OR_EQUAL"]),
        ds=F["D24_UNORM_S8_UINT"], clear_depth=0.6, clear_stencil=126),
    "points": lambda: W.points_scene(400, 256, 160, seed=11),
    "lines": lambda: W.lines_scene(300, 256, 160, seed=12, line_width=3.0),
    "line_strip": lambda: W.lines_scene(200, 256, 160, seed=13, strip=True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_parallel_oracle_is_bit_identical(name):
    assert _same(CASES[name]()) > 0


def test_parallel_oracle_thread_counts():
    scene = W.soup_scene(20_000, 250, 170, seed=7)
    counts = {t: _same(scene, threads=t) for t in (2, 3, 16)}
    assert len(set(counts.values())) == 1
