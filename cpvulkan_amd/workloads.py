"""Synthetic workloads of the draw-path benchmark and their descriptors.

A Scene is one render pass (attachments + one pipeline + draws) held in host
memory.  `materialize` places its buffers in a memory space -- device memory
of a cpvulkan_amd.Context, or host memory -- and returns the C-ABI
descriptors (include/cpv.h) that point at them, so the same scene can be
handed to libcpv and to any other consumer of the same structs.

Generators (all seeded, fp64 then rounded once to fp32):
  cube_scene      15-draw_cube sample: 500x500 B8G8R8A8 + D16, MVP, 36 vertices
                  (Samples/15-draw_cube/15-draw_cube.cpp:38-149,
                  Samples/utils/util_init.cpp:1052-1068, 1541-1693)
  soup_scene      triangle soup, alpha blended, D32 LESS
  grid_scene      textured indexed heightfield with Lambert shading
  skewed_scene    triangle soup with centres from a mixture of Gaussians
  job_scenes      many independent soups of Zipf-distributed sizes
  points_scene    point list with per-vertex gl_PointSize, perspective w
  lines_scene     line list or strip of a given width, perspective w
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class TextureData:
    data: np.ndarray            # uint8, whole mip chain (GetNormalImageSize layout)
    format: int
    width: int
    height: int
    levels: int
    base_level: int = 0
    level_count: int = 0
    sampler: dict = field(default_factory=dict)


@dataclass
class DrawData:
    vertices: dict               # binding -> uint8 array
    count: int
    indices: np.ndarray | None = None
    first: int = 0
    vertex_offset: int = 0
    instance_count: int = 1
    first_instance: int = 0
    viewport: tuple = (0.0, 0.0, 1.0, 1.0, 0.0, 1.0)
    uniforms: dict = field(default_factory=dict)    # binding -> uint8 array
    textures: dict = field(default_factory=dict)    # binding -> TextureData


@dataclass
class Scene:
    name: str
    width: int
    height: int
    color_format: int
    ds_format: int
    pipeline: dict
    draws: list
    color_load: int = abi.LOAD_OP["CLEAR"]
    ds_load: int = abi.LOAD_OP["CLEAR"]
    clear_color: tuple = (0.0, 0.0, 0.0, 0.0)
    clear_depth: float = 1.0
    clear_stencil: int = 0
    initial_color: np.ndarray | None = None
    initial_ds: np.ndarray | None = None
    pipelines: list | None = None        # several pipelines (then draw_pipeline picks one per draw)
    draw_pipeline: list | None = None

    @property
    def color_bytes(self) -> int:
        return self.width * self.height * abi.FORMAT_BYTES[self.color_format] if self.color_format else 0

    @property
    def ds_bytes(self) -> int:
        return self.width * self.height * abi.FORMAT_BYTES[self.ds_format] if self.ds_format else 0

    @property
    def primitive_count(self) -> int:
        topo = self.pipeline.get("topology", abi.TOPOLOGY["TRIANGLE_LIST"])
        T = abi.TOPOLOGY
        n = 0
        for d in self.draws:
            if topo == T["POINT_LIST"]:
                per = d.count
            elif topo == T["LINE_LIST"]:
                per = d.count // 2
            elif topo == T["LINE_STRIP"]:
                per = max(d.count - 1, 0)
            elif topo == T["TRIANGLE_LIST"]:
                per = d.count // 3
            else:
                per = max(d.count - 2, 0)
            n += per * d.instance_count
        return n


def pipeline_desc(p: dict) -> abi.PipelineDesc:
    d = abi.PipelineDesc()
    d.vs_kind = p.get("vs", 0)
    d.fs_kind = p.get("fs", 0)
    d.program = p.get("program", None)
    d.topology = p.get("topology", abi.TOPOLOGY["TRIANGLE_LIST"])
    bindings = p.get("bindings", [])
    attrs = p.get("attributes", [])
    d.binding_count = len(bindings)
    for i, (b, stride, rate) in enumerate(bindings):
        d.bindings[i] = abi.VertexBinding(b, stride, rate, 0)
    d.attribute_count = len(attrs)
    for i, (loc, b, fmt, off) in enumerate(attrs):
        d.attributes[i] = abi.VertexAttribute(loc, b, fmt, off)
    d.cull_mode = p.get("cull", abi.CULL["NONE"])
    d.front_face = p.get("front_face", abi.FRONT_FACE["CCW"])
    d.depth_test = p.get("depth_test", 1)
    d.depth_write = p.get("depth_write", 1)
    d.depth_compare = p.get("depth_compare", abi.COMPARE["LESS"])
    d.depth_bounds_test = p.get("depth_bounds_test", 0)
    d.min_depth_bounds, d.max_depth_bounds = p.get("depth_bounds", (0.0, 1.0))
    d.stencil_test = p.get("stencil_test", 0)
    for side in ("front", "back"):
        st = p.get(side)
        if st:
            setattr(d, side, abi.StencilState(*st, 0))
    d.origin_upper_left = p.get("origin_upper_left", 1)
    d.color_count = 1 if p.get("color_format", 0) else 0
    d.color_formats[0] = p.get("color_format", 0)
    d.depth_stencil_format = p.get("ds_format", 0)
    bl = p.get("blend")
    ba = abi.BlendAttachment()
    ba.color_write_mask = abi.COLOR_RGBA
    if bl:
        ba.blend_enable = 1
        ba.src_color_factor, ba.dst_color_factor, ba.color_op = bl["src_color"], bl["dst_color"], bl["color_op"]
        ba.src_alpha_factor, ba.dst_alpha_factor, ba.alpha_op = bl["src_alpha"], bl["dst_alpha"], bl["alpha_op"]
    d.blend[0] = ba
    for i, v in enumerate(p.get("blend_constants", (0.0, 0.0, 0.0, 0.0))):
        d.blend_constants[i] = v
    d.line_width = p.get("line_width", 1.0)
    return d


class HostMemory:
    """Host buffers addressed by their virtual address (kept alive here)."""

    def __init__(self):
        self.keep: dict[int, np.ndarray] = {}

    def alloc(self, nbytes: int) -> int:
        a = np.zeros(max(int(nbytes), 1), dtype=np.uint8)
        addr = a.ctypes.data
        self.keep[addr] = a
        return addr

    def upload(self, addr: int, data) -> None:
        src = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
        self.keep[addr][: src.size] = src

    def download(self, addr: int, nbytes: int) -> np.ndarray:
        return self.keep[addr][:nbytes].copy()


class DeviceMemory:
    """Device buffers of a cpvulkan_amd.Context."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.addrs: list[int] = []

    def alloc(self, nbytes: int) -> int:
        a = self.ctx.alloc(nbytes)
        self.addrs.append(a)
        return a

    def upload(self, addr: int, data) -> None:
        self.ctx.upload(addr, np.ascontiguousarray(data).view(np.uint8).reshape(-1))

    def download(self, addr: int, nbytes: int) -> np.ndarray:
        return self.ctx.download(addr, nbytes)

    def release(self) -> None:
        for a in self.addrs:
            self.ctx.free(a)
        self.addrs.clear()


@dataclass
class Materialized:
    pass_desc: abi.PassDesc
    pipelines: list
    draws: list
    color_addr: int
    ds_addr: int
    scene: Scene
    mem: object
    draw_pipeline: list | None = None

    def read_color(self) -> np.ndarray:
        return self.mem.download(self.color_addr, self.scene.color_bytes) if self.color_addr else np.zeros(0, np.uint8)

    def read_ds(self) -> np.ndarray:
        return self.mem.download(self.ds_addr, self.scene.ds_bytes) if self.ds_addr else np.zeros(0, np.uint8)

    def reset_attachments(self) -> None:
        """Restore the attachments' initial contents (for LOAD passes)."""
        s = self.scene
        if self.color_addr:
            init = s.initial_color if s.initial_color is not None else np.zeros(s.color_bytes, np.uint8)
            self.mem.upload(self.color_addr, init)
        if self.ds_addr:
            init = s.initial_ds if s.initial_ds is not None else np.zeros(s.ds_bytes, np.uint8)
            self.mem.upload(self.ds_addr, init)


def _sampler(s: dict) -> abi.SamplerDesc:
    out = abi.SamplerDesc()
    out.mag_filter = s.get("mag", abi.FILTER["LINEAR"])
    out.min_filter = s.get("min", abi.FILTER["LINEAR"])
    out.mipmap_mode = s.get("mipmap", 1)
    out.address_u = s.get("address_u", abi.ADDRESS["REPEAT"])
    out.address_v = s.get("address_v", abi.ADDRESS["REPEAT"])
    out.address_w = s.get("address_w", abi.ADDRESS["REPEAT"])
    out.mip_lod_bias = s.get("bias", 0.0)
    out.min_lod = s.get("min_lod", 0.0)
    out.max_lod = s.get("max_lod", 1000.0)
    out.border_color = s.get("border", 0)
    return out


def materialize(scene: Scene, mem) -> Materialized:
    """Copy the scene's buffers into `mem` and build the C-ABI descriptors."""
    pds = [pipeline_desc(dict(p, color_format=scene.color_format, ds_format=scene.ds_format))
           for p in (scene.pipelines or [scene.pipeline])]
    ps = abi.PassDesc()
    ps.width, ps.height = scene.width, scene.height
    color_addr = ds_addr = 0
    if scene.color_format:
        color_addr = mem.alloc(scene.color_bytes)
        ps.color_count = 1
        a = ps.color[0]
        a.address, a.format, a.load_op = color_addr, scene.color_format, scene.color_load
        for i, v in enumerate(scene.clear_color):
            if isinstance(v, float):
                a.clear.f32[i] = v
            else:
                a.clear.u32[i] = int(v) & 0xFFFFFFFF
    if scene.ds_format:
        ds_addr = mem.alloc(scene.ds_bytes)
        ps.has_depth_stencil = 1
        a = ps.depth_stencil
        a.address, a.format, a.load_op = ds_addr, scene.ds_format, scene.ds_load
        a.clear.ds.depth = scene.clear_depth
        a.clear.ds.stencil = scene.clear_stencil
    draws = []
    for d in scene.draws:
        dd = abi.DrawDesc()
        dd.indexed = 1 if d.indices is not None else 0
        dd.count = d.count
        dd.instance_count = d.instance_count
        dd.first = d.first
        dd.vertex_offset = d.vertex_offset
        dd.first_instance = d.first_instance
        if d.indices is not None:
            ia = mem.alloc(d.indices.nbytes)
            mem.upload(ia, d.indices)
            dd.index_address = ia
            dd.index_size = d.indices.dtype.itemsize
        for b, data in d.vertices.items():
            va = mem.alloc(data.nbytes)
            mem.upload(va, data)
            dd.vertex_address[b] = va
            dd.vertex_bytes[b] = data.nbytes
        for i, v in enumerate(d.viewport):
            dd.viewport[i] = v
        for b, data in d.uniforms.items():
            ua = mem.alloc(data.nbytes)
            mem.upload(ua, data)
            dd.uniform_address[b] = ua
        for b, t in d.textures.items():
            ta = mem.alloc(t.data.nbytes)
            mem.upload(ta, t.data)
            td = dd.textures[b]
            td.address, td.format, td.width, td.height = ta, t.format, t.width, t.height
            td.levels, td.base_level = t.levels, t.base_level
            td.level_count = t.level_count or (t.levels - t.base_level)
            td.sampler = _sampler(t.sampler)
        draws.append(dd)
    m = Materialized(ps, pds, draws, color_addr, ds_addr, scene, mem, scene.draw_pipeline)
    m.reset_attachments()
    return m


# ------------------------------------------------------------------ math

def _f32(x):
    return np.float32(x)


def mat4_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Column-major 4x4 product in float32, each column as glm's mat4*vec4
    computes it: (m0*v0 + m1*v1) + (m2*v2 + m3*v3)."""
    out = np.zeros((4, 4), np.float32)   # out[col][row]
    for c in range(4):
        v = b[c]
        add0 = a[0] * v[0] + a[1] * v[1]
        add1 = a[2] * v[2] + a[3] * v[3]
        out[c] = add0 + add1
    return out


def perspective(fovy: float, aspect: float, znear: float, zfar: float) -> np.ndarray:
    """glm::perspective (Samples/utils/glm/gtc/matrix_transform.inl:220-245), float32."""
    t = _f32(math.tan(_f32(fovy) / _f32(2)))
    m = np.zeros((4, 4), np.float32)
    m[0][0] = _f32(1) / (_f32(aspect) * t)
    m[1][1] = _f32(1) / t
    m[2][2] = -(_f32(zfar) + _f32(znear)) / (_f32(zfar) - _f32(znear))
    m[2][3] = _f32(-1)
    m[3][2] = -(_f32(2) * _f32(zfar) * _f32(znear)) / (_f32(zfar) - _f32(znear))
    return m


def look_at(eye, center, up) -> np.ndarray:
    """glm::lookAt (Samples/utils/glm/gtc/matrix_transform.inl:403-428), float32."""
    eye, center, up = (np.asarray(v, np.float32) for v in (eye, center, up))

    def normalize(v):
        return v * (_f32(1) / np.sqrt(np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])))

    def cross(a, b):
        return np.array([a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]], np.float32)

    def dot(a, b):
        return np.float32(a[0] * b[0] + a[1] * b[1] + a[2] * b[2])

    f = normalize(center - eye)
    s = normalize(cross(f, up))
    u = cross(s, f)
    m = np.eye(4, dtype=np.float32)
    m[0][0], m[1][0], m[2][0] = s
    m[0][1], m[1][1], m[2][1] = u
    m[0][2], m[1][2], m[2][2] = -f
    m[3][0] = -dot(s, eye)
    m[3][1] = -dot(u, eye)
    m[3][2] = dot(f, eye)
    return m


def vulkan_clip() -> np.ndarray:
    """Vulkan clip correction of the samples (util_init.cpp:1062): Y inverted, Z halved."""
    return np.array([[1, 0, 0, 0], [0, -1, 0, 0], [0, 0, 0.5, 0], [0, 0, 0.5, 1]], np.float32)


# ------------------------------------------------------------- workloads

# The solid-colour cube of the 15-draw_cube sample: each face is the quad
# (a, b, c, d) drawn as triangles (a, b, c) and (c, b, d).
_CUBE_FACES = [
    (((-1, -1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, 1)), (1, 0, 0)),        # red
    (((-1, -1, -1), (1, -1, -1), (-1, 1, -1), (1, 1, -1)), (0, 1, 0)),    # green
    (((-1, 1, 1), (-1, -1, 1), (-1, 1, -1), (-1, -1, -1)), (0, 0, 1)),    # blue
    (((1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1)), (1, 1, 0)),        # yellow
    (((1, 1, 1), (-1, 1, 1), (1, 1, -1), (-1, 1, -1)), (1, 0, 1)),        # magenta
    (((1, -1, 1), (1, -1, -1), (-1, -1, 1), (-1, -1, -1)), (0, 1, 1)),    # cyan
]


def cube_vertices() -> np.ndarray:
    rows = []
    for (a, b, c, d), col in _CUBE_FACES:
        for p in (a, b, c, c, b, d):
            rows.append([*p, 1.0, *col, 1.0])
    return np.array(rows, np.float32)


def cube_mvp(width: int = 500, height: int = 500) -> np.ndarray:
    fov = _f32(45.0) * _f32(0.01745329251994329576923690768489)
    if width > height:
        fov = fov * (_f32(height) / _f32(width))
    proj = perspective(fov, _f32(width) / _f32(height), 0.1, 100.0)
    view = look_at((-5, 3, -10), (0, 0, 0), (0, -1, 0))
    model = np.eye(4, dtype=np.float32)
    return mat4_mul(mat4_mul(mat4_mul(vulkan_clip(), proj), view), model)


def cube_scene(cull: str = "NONE", width: int = 500, height: int = 500) -> Scene:
    """The 15-draw_cube sample: 500x500 B8G8R8A8 + D16, MVP, 36 vertices."""
    verts = cube_vertices()
    mvp = cube_mvp(width, height)
    pipeline = dict(
        vs=abi.VS["MVP_COLOR"], fs=abi.FS["COLOR"],
        bindings=[(0, 32, 0)],
        attributes=[(0, 0, abi.FORMAT["R32G32B32A32_SFLOAT"], 0), (1, 0, abi.FORMAT["R32G32B32A32_SFLOAT"], 16)],
        cull=abi.CULL[cull], front_face=abi.FRONT_FACE["CW"],
        depth_test=1, depth_write=1, depth_compare=abi.COMPARE["LESS_OR_EQUAL"],
    )
    draw = DrawData(vertices={0: verts.view(np.uint8).reshape(-1)}, count=36,
                    viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0),
                    uniforms={0: mvp.reshape(-1).view(np.uint8)})
    return Scene("cube", width, height, abi.FORMAT["B8G8R8A8_UNORM"], abi.FORMAT["D16_UNORM"], pipeline, [draw],
                 clear_color=(0.2, 0.2, 0.2, 0.2), clear_depth=1.0)


ALPHA_BLEND = dict(src_color=abi.BLEND["SRC_ALPHA"], dst_color=abi.BLEND["ONE_MINUS_SRC_ALPHA"], color_op=0,
                   src_alpha=abi.BLEND["SRC_ALPHA"], dst_alpha=abi.BLEND["ONE_MINUS_SRC_ALPHA"], alpha_op=0)


def soup_triangles(n: int, width: int, height: int, seed: int, centres: np.ndarray | None = None,
                   mu_px: float = math.log(8.0), sigma: float = 1.0) -> np.ndarray:
    """n triangles as 3n vertices of (pos vec4, colour vec4), float32.

    Centres uniform in NDC (or given), edge length L ~ LogNormal(mu, sigma)
    pixels, vertices c + (L/sqrt(3)) (cos t_k, sin t_k) (2/W, 2/H) with
    t_k = t_0 + 2 pi k / 3 + U(-0.3, 0.3), vertices 1 and 2 swapped for a
    random half of the triangles (both windings); z ~ U[0,1]; w = 1;
    rgb ~ U[0,1], a ~ U[0.2, 1] per vertex."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.uniform(-1.0, 1.0, size=(n, 2)) if centres is None else centres
    L = rng.lognormal(mu_px, sigma, size=n)
    t0 = rng.uniform(0.0, 2.0 * math.pi, size=n)
    jit = rng.uniform(-0.3, 0.3, size=(n, 3))
    ang = t0[:, None] + (2.0 * math.pi / 3.0) * np.arange(3)[None, :] + jit
    r = (L / math.sqrt(3.0))[:, None]
    x = c[:, 0:1] + r * np.cos(ang) * (2.0 / width)
    y = c[:, 1:2] + r * np.sin(ang) * (2.0 / height)
    # random winding: half the triangles face each way
    flip = rng.random(n) < 0.5
    x[flip, 1], x[flip, 2] = x[flip, 2].copy(), x[flip, 1].copy()
    y[flip, 1], y[flip, 2] = y[flip, 2].copy(), y[flip, 1].copy()
    z = rng.uniform(0.0, 1.0, size=(n, 3))
    rgb = rng.uniform(0.0, 1.0, size=(n, 3, 3))
    a = rng.uniform(0.2, 1.0, size=(n, 3))
    v = np.empty((n, 3, 8), np.float64)
    v[:, :, 0] = x
    v[:, :, 1] = y
    v[:, :, 2] = z
    v[:, :, 3] = 1.0
    v[:, :, 4:7] = rgb
    v[:, :, 7] = a
    return v.reshape(n * 3, 8).astype(np.float32)


def soup_pipeline() -> dict:
    return dict(
        vs=abi.VS["PASSTHROUGH"], fs=abi.FS["COLOR"],
        bindings=[(0, 32, 0)],
        attributes=[(0, 0, abi.FORMAT["R32G32B32A32_SFLOAT"], 0), (1, 0, abi.FORMAT["R32G32B32A32_SFLOAT"], 16)],
        cull=abi.CULL["NONE"], front_face=abi.FRONT_FACE["CCW"],
        depth_test=1, depth_write=1, depth_compare=abi.COMPARE["LESS"],
        blend=ALPHA_BLEND,
    )


def soup_scene(ntris: int = 1_000_000, width: int = 1920, height: int = 1080, seed: int = 2) -> Scene:
    """Triangle soup: R8G8B8A8 + D32 LESS, alpha blend."""
    verts = soup_triangles(ntris, width, height, seed)
    draw = DrawData(vertices={0: verts.view(np.uint8).reshape(-1)}, count=ntris * 3,
                    viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0))
    return Scene(f"soup{ntris}", width, height, abi.FORMAT["R8G8B8A8_UNORM"], abi.FORMAT["D32_SFLOAT"],
                 soup_pipeline(), [draw], clear_color=(0.0, 0.0, 0.0, 0.0), clear_depth=1.0)


def _clip_positions(ndc: np.ndarray, rng: np.random.Generator, perspective: bool) -> np.ndarray:
    """Clip-space vec4 positions for NDC (x, y) with z ~ U[0, 1]; w ~ U[0.5, 2]
    (positions scaled by w) when perspective, else w = 1."""
    n = ndc.shape[0]
    w = rng.uniform(0.5, 2.0, size=n) if perspective else np.ones(n)
    z = rng.uniform(0.0, 1.0, size=n)
    return np.stack([ndc[:, 0] * w, ndc[:, 1] * w, z * w, w], axis=1)


def points_scene(n: int = 400, width: int = 256, height: int = 160, seed: int = 11, perspective: bool = True,
                 sizes: np.ndarray | None = None) -> Scene:
    """Point list: centres uniform over the viewport and a margin, gl_PointSize
    per vertex (default U[0.5, 14] with a few exact values), colour per
    vertex, alpha blended over D32 LESS.  Vertex layout: pos vec4, colour
    vec4, size float (36 bytes)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ndc = rng.uniform(-1.1, 1.1, size=(n, 2))
    pos = _clip_positions(ndc, rng, perspective)
    if sizes is None:
        sizes = rng.uniform(0.5, 14.0, size=n)
        fixed = np.array([0.0, 1.0, 1.5, 2.0, 3.0, 7.0, 63.96875])
        sizes[: min(n, fixed.size)] = fixed[: min(n, fixed.size)]
    col = np.concatenate([rng.uniform(0.0, 1.0, size=(n, 3)), rng.uniform(0.2, 1.0, size=(n, 1))], axis=1)
    v = np.concatenate([pos, col, np.asarray(sizes, np.float64).reshape(n, 1)], axis=1).astype(np.float32)
    pipe = dict(
        vs=abi.VS["PASSTHROUGH_POINT"], fs=abi.FS["COLOR"], topology=abi.TOPOLOGY["POINT_LIST"],
        bindings=[(0, 36, 0)],
        attributes=[(0, 0, abi.FORMAT["R32G32B32A32_SFLOAT"], 0), (1, 0, abi.FORMAT["R32G32B32A32_SFLOAT"], 16),
                    (2, 0, abi.FORMAT["R32_SFLOAT"], 32)],
        depth_test=1, depth_write=1, depth_compare=abi.COMPARE["LESS"], blend=ALPHA_BLEND,
    )
    draw = DrawData(vertices={0: v.view(np.uint8).reshape(-1)}, count=n,
                    viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0))
    return Scene(f"points{n}", width, height, abi.FORMAT["R8G8B8A8_UNORM"], abi.FORMAT["D32_SFLOAT"], pipe, [draw],
                 clear_color=(0.05, 0.1, 0.15, 1.0), clear_depth=1.0)


def lines_scene(n: int = 300, width: int = 256, height: int = 160, seed: int = 12, strip: bool = False,
                line_width: float = 1.0, perspective: bool = True, length_px: float = 40.0) -> Scene:
    """Line list (2n vertices) or strip (n + 1 vertices) of the given width:
    start points uniform over the viewport and a margin, segments of
    LogNormal(log(length_px), 0.8) pixels in uniform directions, colour per
    vertex, alpha blended over D32 LESS.  Vertex layout: pos vec4, colour
    vec4 (32 bytes)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = n + 1 if strip else 2 * n
    if strip:
        steps = rng.lognormal(math.log(length_px), 0.8, size=m)
        ang = rng.uniform(0.0, 2.0 * math.pi, size=m)
        px = np.cumsum(steps * np.cos(ang)) + rng.uniform(0, width)
        py = np.cumsum(steps * np.sin(ang)) + rng.uniform(0, height)
        ndc = np.stack([(np.mod(px, width * 1.2) - 0.1 * width) * 2.0 / width - 1.0,
                        (np.mod(py, height * 1.2) - 0.1 * height) * 2.0 / height - 1.0], axis=1)
    else:
        a = rng.uniform(-1.1, 1.1, size=(n, 2))
        L = rng.lognormal(math.log(length_px), 0.8, size=n)
        ang = rng.uniform(0.0, 2.0 * math.pi, size=n)
        b = a + np.stack([L * np.cos(ang) * 2.0 / width, L * np.sin(ang) * 2.0 / height], axis=1)
        ndc = np.stack([a, b], axis=1).reshape(m, 2)
    pos = _clip_positions(ndc, rng, perspective)
    col = np.concatenate([rng.uniform(0.0, 1.0, size=(m, 3)), rng.uniform(0.2, 1.0, size=(m, 1))], axis=1)
    v = np.concatenate([pos, col], axis=1).astype(np.float32)
    pipe = soup_pipeline()
    pipe["topology"] = abi.TOPOLOGY["LINE_STRIP" if strip else "LINE_LIST"]
    pipe["line_width"] = line_width
    draw = DrawData(vertices={0: v.view(np.uint8).reshape(-1)}, count=m,
                    viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0))
    return Scene(f"lines{n}", width, height, abi.FORMAT["R8G8B8A8_UNORM"], abi.FORMAT["D32_SFLOAT"], pipe, [draw],
                 clear_color=(0.05, 0.1, 0.15, 1.0), clear_depth=1.0)


def skewed_scene(ntris: int = 4_000_000, width: int = 3840, height: int = 2160, seed: int = 5,
                 clusters: int = 8) -> Scene:
    """Triangle soup whose centres come from a mixture of Gaussians."""
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    means = rng.uniform(-0.8, 0.8, size=(clusters, 2))
    sig = rng.uniform(0.02, 0.2, size=clusters)
    k = rng.integers(0, clusters, size=ntris)
    centres = means[k] + rng.normal(0.0, 1.0, size=(ntris, 2)) * sig[k][:, None]
    verts = soup_triangles(ntris, width, height, seed, centres=centres)
    draw = DrawData(vertices={0: verts.view(np.uint8).reshape(-1)}, count=ntris * 3,
                    viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0))
    return Scene(f"skewed{ntris}", width, height, abi.FORMAT["R8G8B8A8_UNORM"], abi.FORMAT["D32_SFLOAT"],
                 soup_pipeline(), [draw], clear_color=(0.0, 0.0, 0.0, 0.0), clear_depth=1.0)


def _perlin(x: np.ndarray, y: np.ndarray, seed: int) -> np.ndarray:
    """Classic 2-D gradient (Perlin) noise, fp64, values in about [-0.7, 0.7]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    perm = rng.permutation(256)
    perm = np.concatenate([perm, perm])
    ang = rng.uniform(0.0, 2.0 * math.pi, size=256)
    grad = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    xi, yi = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64)
    xf, yf = x - xi, y - yi
    xi &= 255
    yi &= 255

    def g(ix, iy, dx, dy):
        h = perm[perm[ix] + iy]
        return grad[h, 0] * dx + grad[h, 1] * dy

    def fade(t):
        return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)

    u, v = fade(xf), fade(yf)
    n00 = g(xi, yi, xf, yf)
    n10 = g(xi + 1, yi, xf - 1.0, yf)
    n01 = g(xi, yi + 1, xf, yf - 1.0)
    n11 = g(xi + 1, yi + 1, xf - 1.0, yf - 1.0)
    a = n00 + u * (n10 - n00)
    b = n01 + u * (n11 - n01)
    return a + v * (b - a)


def noise_texture(size: int = 1024, seed: int = 7) -> TextureData:
    """size^2 RGBA8 texture of uniform U8 noise with its full mip chain, each
    level the rounded 2x2 box filter of the one above, laid out level after
    level as GetNormalImageSize places them (CPVulkanBase/Formats.cpp:455-482)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lv = rng.integers(0, 256, size=(size, size, 4), dtype=np.uint16)
    levels = [lv.astype(np.uint8)]
    while lv.shape[0] > 1 or lv.shape[1] > 1:
        h, w = max(lv.shape[0] // 2, 1), max(lv.shape[1] // 2, 1)
        src = lv
        if src.shape[0] == 1:
            src = np.concatenate([src, src], axis=0)
        if src.shape[1] == 1:
            src = np.concatenate([src, src], axis=1)
        s = src[0::2, 0::2] + src[1::2, 0::2] + src[0::2, 1::2] + src[1::2, 1::2]
        lv = ((s + 2) // 4)[:h, :w]
        levels.append(lv.astype(np.uint8))
    data = np.concatenate([l.reshape(-1) for l in levels])
    return TextureData(data=data, format=abi.FORMAT["R8G8B8A8_UNORM"], width=size, height=size, levels=len(levels),
                       sampler=dict(mag=abi.FILTER["LINEAR"], min=abi.FILTER["LINEAR"], mipmap=1,
                                    address_u=abi.ADDRESS["REPEAT"], address_v=abi.ADDRESS["REPEAT"],
                                    min_lod=0.0, max_lod=11.0, bias=0.0))


def grid_mvp(width: int, height: int, orbit_deg: float = 34.0) -> np.ndarray:
    """Vulkan clip * perspective(60 deg) * lookAt from an orbit around the grid."""
    fov = _f32(60.0) * _f32(0.01745329251994329576923690768489)
    proj = perspective(fov, _f32(width) / _f32(height), 0.1, 100.0)
    th = math.radians(orbit_deg)
    eye = (0.72 * math.sin(th), 0.8, -0.72 * math.cos(th))
    view = look_at(eye, (0.0, -0.1, 0.0), (0.0, 1.0, 0.0))
    return mat4_mul(mat4_mul(vulkan_clip(), proj), view)


def grid_scene(n: int = 512, width: int = 3840, height: int = 2160, seed: int = 3, tex_size: int = 1024,
               uv_repeat: float = 4.0, orbit_deg: float = 34.0) -> Scene:
    """Textured indexed heightfield with Lambert shading.

    n x n vertices over [-1,1]^2 in xz with y = 0.2 * Perlin; vertex = pos vec3,
    normal vec3, uv vec2 (32 B); uint32 indices, two triangles (a, b, c),
    (c, b, d) per cell; texture tex_size^2 RGBA8 noise with a full mip chain,
    LINEAR/LINEAR/MIPMAP_LINEAR, REPEAT; R8G8B8A8 + D32 LESS, no blend.
    Fragment colour = texture(uv) * max(dot(normalize(n), L), 0)."""
    lin = np.linspace(-1.0, 1.0, n)
    X, Z = np.meshgrid(lin, lin)                       # row = z, column = x
    f = 3.0
    H = 0.2 * _perlin((X + 1.0) * f, (Z + 1.0) * f, seed)
    eps = 1e-3
    dhx = 0.2 * (_perlin((X + 1.0 + eps) * f, (Z + 1.0) * f, seed) - _perlin((X + 1.0 - eps) * f, (Z + 1.0) * f, seed)) / (2 * eps)
    dhz = 0.2 * (_perlin((X + 1.0) * f, (Z + 1.0 + eps) * f, seed) - _perlin((X + 1.0) * f, (Z + 1.0 - eps) * f, seed)) / (2 * eps)
    N = np.stack([-dhx, np.ones_like(H), -dhz], axis=-1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    v = np.empty((n, n, 8), np.float64)
    v[..., 0], v[..., 1], v[..., 2] = X, H, Z
    v[..., 3:6] = N
    v[..., 6] = (X + 1.0) * 0.5 * uv_repeat
    v[..., 7] = (Z + 1.0) * 0.5 * uv_repeat
    verts = v.reshape(-1, 8).astype(np.float32)
    i = np.arange(n - 1)
    a = (i[:, None] * n + i[None, :]).reshape(-1)
    b, c = a + 1, a + n
    d = c + 1
    idx = np.stack([a, b, c, c, b, d], axis=1).reshape(-1).astype(np.uint32)
    light = np.array([0.3, 0.9, -0.2], np.float64)
    light = np.concatenate([light / np.linalg.norm(light), [0.0]]).astype(np.float32)
    pipeline = dict(
        vs=abi.VS["MVP_LAMBERT"], fs=abi.FS["TEX_LAMBERT"],
        bindings=[(0, 32, 0)],
        attributes=[(0, 0, abi.FORMAT["R32G32B32_SFLOAT"], 0), (1, 0, abi.FORMAT["R32G32B32_SFLOAT"], 12),
                    (2, 0, abi.FORMAT["R32G32_SFLOAT"], 24)],
        cull=abi.CULL["NONE"], front_face=abi.FRONT_FACE["CCW"],
        depth_test=1, depth_write=1, depth_compare=abi.COMPARE["LESS"],
    )
    draw = DrawData(vertices={0: verts.view(np.uint8).reshape(-1)}, count=idx.size, indices=idx,
                    viewport=(0.0, 0.0, float(width), float(height), 0.0, 1.0),
                    uniforms={0: grid_mvp(width, height, orbit_deg).reshape(-1).view(np.uint8),
                              2: light.view(np.uint8)},
                    textures={1: noise_texture(tex_size, seed + 100)})
    return Scene(f"grid{n}", width, height, abi.FORMAT["R8G8B8A8_UNORM"], abi.FORMAT["D32_SFLOAT"], pipeline, [draw],
                 clear_color=(0.0, 0.0, 0.0, 0.0), clear_depth=1.0)


def zipf_counts(n: int, rng: np.random.Generator, s: float = 1.2, lo: int = 12, hi: int = 200_000) -> np.ndarray:
    """Triangle counts ~ Zipf(s) truncated to [lo, hi], by inverse CDF of the
    continuous power law on [lo, hi + 1), rounded down."""
    u = rng.random(n)
    a, b = lo ** (1.0 - s), (hi + 1) ** (1.0 - s)
    k = np.floor((a + u * (b - a)) ** (1.0 / (1.0 - s))).astype(np.int64)
    return np.clip(k, lo, hi)


def job_scenes(n_jobs: int = 256, seed: int = 4, sides=(256, 512, 1024, 2048, 4096), lo: int = 12,
               hi: int = 200_000) -> list:
    """Independent render jobs of uneven size.  Triangle
    counts ~ Zipf(1.2) over [lo, hi]; framebuffer side uniform over `sides`
    (square); R8G8B8A8 + D32; triangles as soup_scene, from per-job seeds, in
    each job's own pixel units."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = zipf_counts(n_jobs, rng, lo=lo, hi=hi)
    side = rng.choice(np.asarray(sides), size=n_jobs)
    jobs = []
    for j in range(n_jobs):
        s = soup_scene(int(counts[j]), int(side[j]), int(side[j]), seed=seed * 1_000_003 + j)
        s.name = f"job{j}-{int(side[j])}-{int(counts[j])}"
        jobs.append(s)
    return jobs


def job_desc(m: Materialized) -> abi.JobDesc:
    """The C-ABI job (one pass, one pipeline, one draw) of a materialised scene."""
    if len(m.draws) != 1 or len(m.pipelines) != 1:
        raise ValueError("a job holds one pipeline and one draw")
    j = abi.JobDesc()
    j.pass_ = m.pass_desc
    j.pipeline = m.pipelines[0]
    j.draw = m.draws[0]
    return j
