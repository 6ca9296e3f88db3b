"""SPIR-V -> HIP translation on the CPU: the generated source for the
benchmark shaders, hiprtc compilation of every pair (from the program
cache after the first build), and the errors unsupported modules raise."""
import numpy as np
import pytest

import cpvulkan_amd
from cpvulkan_amd import spirv as S


def test_translation_of_the_cube_shaders():
    src = cpvulkan_amd.spirv_translate(S.vs_mvp_color(), S.fs_color())
    # gl_Position = mvp * pos through glm's mat4 * vec4 order; the colour varying is 4 components
    assert "mat4_mul_vec4(" in src and "env.uniform[0]" in src
    assert "static constexpr int kOutputs = 4;" in src and "static constexpr int kInputs = 4;" in src
    assert "kMayDiscard = false" in src
    # 7 tile configurations, points/lines, the 7 again for the fixed formats
    assert src.count("template __global__ void cpv::raster_kernel<cpv::SpvFs") == 15


def test_translation_of_point_size():
    """gl_PointSize travels as the vertex output after the varyings only
    when the shader writes it."""
    src = cpvulkan_amd.spirv_translate(S.vs_point_size(), S.fs_color())
    assert "static constexpr int kOutputs = 5;" in src and "out[4] = " in src
    assert "cpv::raster_kernel<cpv::SpvFs, 16, 4, true>" in src


def test_translation_of_the_lambert_shaders():
    src = cpvulkan_amd.spirv_translate(S.vs_mvp_lambert(), S.fs_tex_lambert())
    assert "fetch_attribute(env, 0, 3," in src and "fetch_attribute(env, 2, 2," in src
    assert "sample_implicit(" in src and "&env.tex[1]" in src
    assert "1.0f / sqrtf(" in src                # glm::normalize: x * (1 / sqrt(x.x*x.x + ...))
    assert "spv_max(" in src                      # std::max


def test_translation_of_discard():
    src = cpvulkan_amd.spirv_translate(S.vs_passthrough(), S.fs_discard_dark())
    assert "kMayDiscard = true" in src and "return true;" in src
    assert "if (" in src and "goto L" in src


@pytest.mark.parametrize("vs,fs", [("vs_mvp_color", "fs_color"), ("vs_passthrough", "fs_color"),
                                   ("vs_mvp_lambert", "fs_tex_lambert"), ("vs_passthrough", "fs_discard_dark")])
def test_hiprtc_builds(vs, fs):
    assert cpvulkan_amd.spirv_build(S.SHADERS[vs](), S.SHADERS[fs]()) > 10_000


def _with_call():
    """A module whose entry point calls a function that does not exist."""
    m = S.Module()
    v4 = m.vec(4)
    out = m.variable("Output", v4, "o")
    m.decorate(out, "Location", 0)
    f = S.Function(m)
    f.op("FunctionCall", m.void(), 999)
    return m.assemble("Fragment", f.end(), [out], True)


def test_unsupported_constructs_are_reported():
    with pytest.raises(cpvulkan_amd.CpvError, match="call of an undefined function"):
        cpvulkan_amd.spirv_translate(S.vs_passthrough(), _with_call())
    with pytest.raises(cpvulkan_amd.CpvError, match="magic"):
        cpvulkan_amd.spirv_translate([1, 2, 3, 4, 5], S.fs_color())
    with pytest.raises(cpvulkan_amd.CpvError, match="not a fragment shader"):
        cpvulkan_amd.spirv_translate(S.vs_passthrough(), S.vs_passthrough())
    with pytest.raises(cpvulkan_amd.CpvError, match="entry point 'other' not found"):
        cpvulkan_amd.spirv_translate(S.vs_passthrough(), S.fs_color(), fs_entry="other")


def test_fragment_input_without_vertex_output():
    """fs_tex_lambert reads locations 0 (vec3) and 1 (vec2); the passthrough
    vertex shader writes only location 0 as vec4, so location 1 is missing."""
    with pytest.raises(cpvulkan_amd.CpvError, match="location 1 has no vertex output"):
        cpvulkan_amd.spirv_translate(S.vs_passthrough(), S.fs_tex_lambert())
