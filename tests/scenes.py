"""Scene variants for the parity tests: small render passes that exercise one
piece of fixed-function state each (compare ops, blend, formats, stencil,
topologies, indexed and instanced draws, load ops, multi-draw passes)."""
from __future__ import annotations

import copy
import math

import numpy as np

from cpvulkan_amd import abi
from cpvulkan_amd import workloads as W
from cpvulkan_amd.workloads import DrawData, Scene

F = abi.FORMAT


def soup(n=3000, width=160, height=120, seed=0, w_range=None, size_px=math.log(8.0), offscreen=0.0) -> Scene:
    """Triangle soup; w_range=(lo, hi) gives every vertex a random clip w
    (x, y, z scaled by it, so NDC positions are unchanged)."""
    s = W.soup_scene(n, width, height, seed)
    if w_range is not None or size_px != math.log(8.0) or offscreen:
        rng = np.random.Generator(np.random.PCG64(seed + 77))
        centres = rng.uniform(-1.0 - offscreen, 1.0 + offscreen, size=(n, 2))
        v = W.soup_triangles(n, width, height, seed, centres=centres, mu_px=size_px)
        if w_range is not None:
            w = rng.uniform(w_range[0], w_range[1], size=v.shape[0]).astype(np.float32)
            v[:, 0:3] *= w[:, None]
            v[:, 3] = w
        s.draws[0].vertices[0] = v.view(np.uint8).reshape(-1)
    return s


def with_pipeline(s: Scene, **kw) -> Scene:
    s = copy.deepcopy(s)
    s.pipeline.update(kw)
    return s


def with_formats(s: Scene, color=None, ds=None, clear_color=None, clear_depth=None, clear_stencil=None) -> Scene:
    s = copy.deepcopy(s)
    if color is not None:
        s.color_format = color
    if ds is not None:
        s.ds_format = ds
    if clear_color is not None:
        s.clear_color = clear_color
    if clear_depth is not None:
        s.clear_depth = clear_depth
    if clear_stencil is not None:
        s.clear_stencil = clear_stencil
    return s


def with_viewport(s: Scene, vp) -> Scene:
    s = copy.deepcopy(s)
    for d in s.draws:
        d.viewport = vp
    return s


def indexed(s: Scene, index_dtype=np.uint32, vertex_offset=0, topology="TRIANGLE_LIST", seed=0) -> Scene:
    """Re-express the soup through an index buffer over a shuffled vertex
    buffer (optionally with a vertex offset)."""
    s = copy.deepcopy(s)
    d = s.draws[0]
    verts = d.vertices[0].view(np.float32).reshape(-1, 8)
    nv = verts.shape[0]
    if index_dtype == np.uint16 and nv + vertex_offset > 65535:
        verts = verts[: 65535 - vertex_offset - 1]
        nv = verts.shape[0] - verts.shape[0] % 3
        verts = verts[:nv]
    rng = np.random.Generator(np.random.PCG64(seed))
    perm = rng.permutation(nv)
    buf = np.zeros((nv + vertex_offset, 8), np.float32)
    buf[vertex_offset + perm] = verts
    idx = perm.astype(index_dtype)
    d.vertices[0] = buf.view(np.uint8).reshape(-1)
    d.indices = idx
    d.count = nv
    d.vertex_offset = vertex_offset
    s.pipeline["topology"] = abi.TOPOLOGY[topology]
    return s


def strip(n=400, width=160, height=120, seed=0, fan=False) -> Scene:
    """A random walk strip (or fan) of n+2 vertices."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pts = np.cumsum(rng.normal(0.0, 0.06, size=(n + 2, 2)), axis=0)
    pts = np.clip(pts, -1.2, 1.2)
    if fan:
        pts[0] = (0.0, 0.0)
    v = np.zeros((n + 2, 8), np.float32)
    v[:, 0:2] = pts
    v[:, 2] = rng.uniform(0.0, 1.0, n + 2)
    v[:, 3] = 1.0
    v[:, 4:8] = rng.uniform(0.2, 1.0, size=(n + 2, 4))
    s = W.soup_scene(1, width, height, seed)
    s.draws[0].vertices[0] = v.view(np.uint8).reshape(-1)
    s.draws[0].count = n + 2
    s.pipeline["topology"] = abi.TOPOLOGY["TRIANGLE_FAN" if fan else "TRIANGLE_STRIP"]
    return s


def instanced(n=500, width=160, height=120, seed=0, instances=3) -> Scene:
    """Per-instance colour from an instance-rate binding; the positions of
    each instance are offset by... nothing (instances overdraw in order)."""
    s = soup(n, width, height, seed)
    d = s.draws[0]
    verts = d.vertices[0].view(np.float32).reshape(-1, 8)
    pos = np.ascontiguousarray(verts[:, 0:4])
    rng = np.random.Generator(np.random.PCG64(seed + 5))
    inst = rng.uniform(0.0, 1.0, size=(instances + 2, 4)).astype(np.float32)
    d.vertices = {0: pos.view(np.uint8).reshape(-1), 1: inst.view(np.uint8).reshape(-1)}
    d.instance_count = instances
    d.first_instance = 1
    s.pipeline["bindings"] = [(0, 16, 0), (1, 16, 1)]
    s.pipeline["attributes"] = [(0, 0, F["R32G32B32A32_SFLOAT"], 0), (1, 1, F["R32G32B32A32_SFLOAT"], 0)]
    s.pipeline["depth_compare"] = abi.COMPARE["LESS_OR_EQUAL"]
    return s


def unorm_colour_attribute(n=800, seed=0) -> Scene:
    """Colour attribute read from an R8G8B8A8_UNORM vertex format (GetPixel path)."""
    s = soup(n, 160, 120, seed)
    d = s.draws[0]
    verts = d.vertices[0].view(np.float32).reshape(-1, 8)
    rng = np.random.Generator(np.random.PCG64(seed + 9))
    out = np.zeros(verts.shape[0], dtype=[("p", np.float32, 4), ("c", np.uint8, 4), ("pad", np.uint8, 12)])
    out["p"] = verts[:, 0:4]
    out["c"] = rng.integers(0, 256, size=(verts.shape[0], 4), dtype=np.uint8)
    d.vertices[0] = out.view(np.uint8).reshape(-1)
    s.pipeline["attributes"] = [(0, 0, F["R32G32B32A32_SFLOAT"], 0), (1, 0, F["R8G8B8A8_UNORM"], 16)]
    return s


def load_op(s: Scene, seed=0, dont_care=False) -> Scene:
    s = copy.deepcopy(s)
    rng = np.random.Generator(np.random.PCG64(seed))
    op = abi.LOAD_OP["DONT_CARE" if dont_care else "LOAD"]
    s.color_load = op
    s.ds_load = op
    s.initial_color = rng.integers(0, 256, size=s.color_bytes, dtype=np.uint8)
    if s.ds_format == F["D32_SFLOAT"]:
        s.initial_ds = rng.uniform(0.3, 1.0, size=s.width * s.height).astype(np.float32).view(np.uint8)
    else:
        s.initial_ds = rng.integers(0, 256, size=s.ds_bytes, dtype=np.uint8)
    return s


def multi_draw(seed=0) -> Scene:
    """Three draws: soup A (blend), soup B (different pipeline: no blend,
    LEQUAL), soup C (pipeline A again)."""
    a = soup(1500, 160, 120, seed)
    b = soup(1500, 160, 120, seed + 1)
    c = soup(1500, 160, 120, seed + 2)
    s = copy.deepcopy(a)
    s.draws = [a.draws[0], b.draws[0], c.draws[0]]
    pa = copy.deepcopy(a.pipeline)
    pb = copy.deepcopy(a.pipeline)
    pb.pop("blend", None)
    pb["depth_compare"] = abi.COMPARE["LESS_OR_EQUAL"]
    s.pipelines = [pa, pb]
    s.draw_pipeline = [0, 1, 0]
    return s
