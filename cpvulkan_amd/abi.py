"""ctypes mirror of the C-ABI structs in include/cpv.h.

The layouts are checked against the compiled library at load time
(cpv_abi_sizes), so a drift between this file and the header fails loudly.
"""
from __future__ import annotations

import ctypes as C

MAX_VERTEX_BINDINGS = 16
MAX_VERTEX_ATTRIBUTES = 16
MAX_COLOR_ATTACHMENTS = 1
MAX_UNIFORMS = 4
MAX_TEXTURES = 4

# VkFormat values (include/cpv_vk_enums.h)
FORMAT = dict(
    UNDEFINED=0, R8_UNORM=9, R8G8B8A8_UNORM=37, R8G8B8A8_SNORM=38, R8G8B8A8_UINT=41, R8G8B8A8_SINT=42,
    B8G8R8A8_UNORM=44, R16G16B16A16_UNORM=91, R32_SFLOAT=100, R32G32_SFLOAT=103, R32G32B32_SFLOAT=106,
    R32G32B32A32_SFLOAT=109, D16_UNORM=124, X8_D24_UNORM_PACK32=125, D32_SFLOAT=126, S8_UINT=127,
    D16_UNORM_S8_UINT=128, D24_UNORM_S8_UINT=129, D32_SFLOAT_S8_UINT=130,
    R16_UNORM=70, R16_SINT=75, R16G16_UINT=81, R16G16_SINT=82, R32_UINT=98, R32_SINT=99, R32G32_UINT=101,
    R32G32B32A32_UINT=107, R32G32B32A32_SINT=108,
)
COMPARE = dict(NEVER=0, LESS=1, EQUAL=2, LESS_OR_EQUAL=3, GREATER=4, NOT_EQUAL=5, GREATER_OR_EQUAL=6, ALWAYS=7)
STENCIL = dict(KEEP=0, ZERO=1, REPLACE=2, INC_CLAMP=3, DEC_CLAMP=4, INVERT=5, INC_WRAP=6, DEC_WRAP=7)
BLEND = dict(ZERO=0, ONE=1, SRC_COLOR=2, ONE_MINUS_SRC_COLOR=3, DST_COLOR=4, ONE_MINUS_DST_COLOR=5, SRC_ALPHA=6,
             ONE_MINUS_SRC_ALPHA=7, DST_ALPHA=8, ONE_MINUS_DST_ALPHA=9, CONSTANT_COLOR=10,
             ONE_MINUS_CONSTANT_COLOR=11, CONSTANT_ALPHA=12, ONE_MINUS_CONSTANT_ALPHA=13, SRC_ALPHA_SATURATE=14)
BLEND_OP = dict(ADD=0, SUBTRACT=1, REVERSE_SUBTRACT=2, MIN=3, MAX=4)
CULL = dict(NONE=0, FRONT=1, BACK=2, FRONT_AND_BACK=3)
FRONT_FACE = dict(CCW=0, CW=1)
TOPOLOGY = dict(POINT_LIST=0, LINE_LIST=1, LINE_STRIP=2, TRIANGLE_LIST=3, TRIANGLE_STRIP=4, TRIANGLE_FAN=5)
LOAD_OP = dict(LOAD=0, CLEAR=1, DONT_CARE=2)
FILTER = dict(NEAREST=0, LINEAR=1)
ADDRESS = dict(REPEAT=0, MIRRORED_REPEAT=1, CLAMP_TO_EDGE=2, CLAMP_TO_BORDER=3, MIRROR_CLAMP_TO_EDGE=4)
VS = dict(PASSTHROUGH=1, MVP_COLOR=2, MVP_LAMBERT=3, PASSTHROUGH_POINT=4)
FS = dict(COLOR=1, TEX_LAMBERT=2, COLOR_DISCARD=3, COLOR_FLAT=4, COLOR_NOPERSPECTIVE=5)
SCHED = dict(GRID=0, COUNTER=1, PREDICTED=2)
# non-error status codes (fences)
OK, NOT_READY, TIMEOUT = 0, 1, 2
COLOR_RGBA = 15

FORMAT_BYTES = {9: 1, 37: 4, 38: 4, 41: 4, 42: 4, 44: 4, 91: 8, 100: 4, 103: 8, 106: 12, 109: 16,
                124: 2, 125: 4, 126: 4, 127: 1, 128: 3, 129: 4, 130: 8,
                70: 2, 75: 2, 81: 4, 82: 4, 98: 4, 99: 4, 101: 8, 107: 16, 108: 16}

u32 = C.c_uint32
i32 = C.c_int32
u64 = C.c_uint64
f32 = C.c_float


class VertexBinding(C.Structure):
    _fields_ = [("binding", u32), ("stride", u32), ("input_rate", u32), ("reserved", u32)]


class VertexAttribute(C.Structure):
    _fields_ = [("location", u32), ("binding", u32), ("format", u32), ("offset", u32)]


class StencilState(C.Structure):
    _fields_ = [("fail_op", u32), ("pass_op", u32), ("depth_fail_op", u32), ("compare_op", u32),
                ("compare_mask", u32), ("write_mask", u32), ("reference", u32), ("reserved", u32)]


class BlendAttachment(C.Structure):
    _fields_ = [("blend_enable", u32), ("src_color_factor", u32), ("dst_color_factor", u32), ("color_op", u32),
                ("src_alpha_factor", u32), ("dst_alpha_factor", u32), ("alpha_op", u32), ("color_write_mask", u32)]


class PipelineDesc(C.Structure):
    _fields_ = [
        ("vs_kind", u32), ("fs_kind", u32), ("program", C.c_void_p),
        ("topology", u32), ("primitive_restart", u32), ("binding_count", u32), ("attribute_count", u32),
        ("bindings", VertexBinding * MAX_VERTEX_BINDINGS),
        ("attributes", VertexAttribute * MAX_VERTEX_ATTRIBUTES),
        ("cull_mode", u32), ("front_face", u32), ("polygon_mode", u32), ("rasterizer_discard", u32),
        ("depth_clamp", u32), ("depth_test", u32), ("depth_write", u32), ("depth_compare", u32),
        ("depth_bounds_test", u32), ("stencil_test", u32), ("logic_op_enable", u32), ("origin_upper_left", u32),
        ("min_depth_bounds", f32), ("max_depth_bounds", f32),
        ("front", StencilState), ("back", StencilState),
        ("blend_constants", f32 * 4),
        ("color_count", u32),
        ("color_formats", u32 * MAX_COLOR_ATTACHMENTS),
        ("depth_stencil_format", u32),
        ("blend", BlendAttachment * MAX_COLOR_ATTACHMENTS),
        ("line_width", f32), ("reserved", u32),
    ]


class SamplerDesc(C.Structure):
    _fields_ = [("mag_filter", u32), ("min_filter", u32), ("mipmap_mode", u32),
                ("address_u", u32), ("address_v", u32), ("address_w", u32),
                ("mip_lod_bias", f32), ("min_lod", f32), ("max_lod", f32),
                ("border_color", u32), ("unnormalized", u32), ("reserved", u32)]


class TextureDesc(C.Structure):
    _fields_ = [("address", u64), ("format", u32), ("width", u32), ("height", u32), ("levels", u32),
                ("base_level", u32), ("level_count", u32), ("sampler", SamplerDesc)]


class DrawDesc(C.Structure):
    _fields_ = [
        ("indexed", u32), ("count", u32), ("instance_count", u32), ("first", u32),
        ("vertex_offset", i32), ("first_instance", u32), ("index_size", u32), ("reserved", u32),
        ("index_address", u64),
        ("vertex_address", u64 * MAX_VERTEX_BINDINGS),
        ("vertex_bytes", u64 * MAX_VERTEX_BINDINGS),
        ("viewport", f32 * 6),
        ("reserved2", u32 * 2),
        ("uniform_address", u64 * MAX_UNIFORMS),
        ("textures", TextureDesc * MAX_TEXTURES),
    ]


class _DS(C.Structure):
    _fields_ = [("depth", f32), ("stencil", u32)]


class _Clear(C.Union):
    _fields_ = [("f32", f32 * 4), ("u32", u32 * 4), ("i32", i32 * 4), ("ds", _DS)]


class AttachmentDesc(C.Structure):
    _fields_ = [("address", u64), ("format", u32), ("load_op", u32), ("clear", _Clear)]


class PassDesc(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("color_count", u32), ("has_depth_stencil", u32),
                ("color", AttachmentDesc * MAX_COLOR_ATTACHMENTS), ("depth_stencil", AttachmentDesc)]


class JobDesc(C.Structure):
    _fields_ = [("pass_", PassDesc), ("pipeline", PipelineDesc), ("draw", DrawDesc)]


class Config(C.Structure):
    _fields_ = [("tile_size", u32), ("waves_per_tile", u32), ("scheduler", u32), ("split_threshold", u32),
                ("record_costs", u32), ("count_fragments", u32), ("reserved", u32 * 2)]


class TileTrace(C.Structure):
    _fields_ = [("features", f32 * 12), ("cycles", f32), ("job", u32), ("tile", u32), ("reserved", u32)]


class AutotuneEntry(C.Structure):
    _fields_ = [("tile_size", u32), ("waves_per_tile", u32), ("ms", f32), ("reserved", u32)]


class CopyRegion(C.Structure):
    _fields_ = [("src", u64), ("dst", u64), ("src_pitch", u64), ("dst_pitch", u64), ("row_bytes", u64), ("rows", u32),
                ("pad", u32)]


class ImageLevel(C.Structure):
    _fields_ = [("address", u64), ("format", u32), ("width", u32), ("height", u32), ("pad", u32)]


class Rect(C.Structure):
    _fields_ = [("x", i32), ("y", i32), ("width", u32), ("height", u32)]


class BlitRegion(C.Structure):
    _fields_ = [("src", i32 * 4), ("dst", i32 * 4), ("aspect", u32), ("pad", u32)]


MLP_FEATURES = 12
MLP_HIDDEN = 64
MLP_PARAMS = 12 * 64 + 64 + 64 * 64 + 64 + 64 + 1

STRUCT_ORDER = [PipelineDesc, DrawDesc, PassDesc, AttachmentDesc, TextureDesc, SamplerDesc, JobDesc, Config, TileTrace,
                AutotuneEntry, CopyRegion, ImageLevel, Rect, BlitRegion]


def check_layout(lib: C.CDLL) -> None:
    """Compare ctypes sizes with the library's sizeof() values."""
    out = (u32 * 16)()
    n = lib.cpv_abi_sizes(out, 16)
    if n != len(STRUCT_ORDER):
        raise RuntimeError(f"cpv_abi_sizes reports {n} structs, binding knows {len(STRUCT_ORDER)}")
    for i, cls in enumerate(STRUCT_ORDER):
        if out[i] != C.sizeof(cls):
            raise RuntimeError(f"ABI layout mismatch for {cls.__name__}: C {out[i]} vs ctypes {C.sizeof(cls)}")
