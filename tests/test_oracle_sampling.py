"""Pins the oracle's texture sampling (ImageSampler.cpp:12-673,
GlslFunctions.cpp:598-671) with properties the mathematics fixes:

* bilinear filtering of texels that are an affine function of the texel
  coordinates reproduces that function at every interior sample point (a
  swapped u/v, a missing -0.5 texel-centre offset or a wrong lerp order
  breaks it);
* texel centres return the texel itself;
* the wrap modes against their definitions on small n (REPEAT (v%n+n)%n,
  MIRRORED_REPEAT, CLAMP_TO_EDGE, CLAMP_TO_BORDER with the border table);
* implicit LOD is level 0 (adding or corrupting other mip levels changes
  nothing), the LOD bias selects levels as SampleImage does;
* a 2x2 texture bilinear at one point against a hand evaluation.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from cpvulkan_amd import abi

F = abi.FORMAT
A = abi.ADDRESS
FL = abi.FILTER


def make_texture(levels_data, fmt=F["R32G32B32A32_SFLOAT"], **sampler):
    """levels_data: list of (h, w, 4) arrays, one per mip level."""
    h0, w0 = levels_data[0].shape[:2]
    dt = np.float32 if fmt == F["R32G32B32A32_SFLOAT"] else np.uint8
    blob = np.concatenate([np.ascontiguousarray(l, dtype=dt).reshape(-1).view(np.uint8) for l in levels_data])
    t = abi.TextureDesc()
    t.address = blob.ctypes.data
    t.format = fmt
    t.width, t.height = w0, h0
    t.levels = len(levels_data)
    t.base_level = 0
    t.level_count = len(levels_data)
    s = t.sampler
    s.mag_filter = sampler.get("mag", FL["LINEAR"])
    s.min_filter = sampler.get("min", FL["LINEAR"])
    s.mipmap_mode = sampler.get("mipmap", 1)
    s.address_u = sampler.get("address_u", A["REPEAT"])
    s.address_v = sampler.get("address_v", A["REPEAT"])
    s.address_w = A["REPEAT"]
    s.mip_lod_bias = sampler.get("bias", 0.0)
    s.min_lod = sampler.get("min_lod", 0.0)
    s.max_lod = sampler.get("max_lod", 1000.0)
    s.border_color = sampler.get("border", 0)
    return t, blob


def sample(tex, u, v):
    out = (C.c_float * 4)()
    assert oracle_lib.lib().cpvref_sample_implicit(C.byref(tex), u, v, out) == 0
    return np.array(out, np.float32)


def mips_of(level0):
    lv = [level0]
    while lv[-1].shape[0] > 1 or lv[-1].shape[1] > 1:
        h, w = max(lv[-1].shape[0] // 2, 1), max(lv[-1].shape[1] // 2, 1)
        lv.append(np.zeros((h, w, 4), np.float32))
    return lv


def test_bilinear_reproduces_affine_data():
    W, H = 16, 8
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    level = np.zeros((H, W, 4), np.float32)
    level[..., 0] = 3.0 * xs + 0.25     # distinct x / y slopes so a transpose shows
    level[..., 1] = -5.0 * ys + 1.0
    level[..., 2] = xs + 7.0 * ys
    level[..., 3] = 1.0
    tex, keep = make_texture([level], address_u=A["CLAMP_TO_EDGE"], address_v=A["CLAMP_TO_EDGE"])
    rng = np.random.default_rng(0)
    for _ in range(400):
        # interior: texel-centre coordinates x in [0, W-1], y in [0, H-1]
        x, y = rng.uniform(0, W - 1), rng.uniform(0, H - 1)
        u, v = np.float32((x + 0.5) / W), np.float32((y + 0.5) / H)
        xe = float(u) * W - 0.5      # the sample point the float coordinates name
        ye = float(v) * H - 0.5
        got = sample(tex, u, v)
        want = np.array([3.0 * xe + 0.25, -5.0 * ye + 1.0, xe + 7.0 * ye, 1.0])
        assert np.allclose(got, want, rtol=0, atol=2e-4), (x, y, got, want)


def test_texel_centres_return_texels():
    rng = np.random.default_rng(1)
    level = rng.uniform(-4, 4, size=(4, 8, 4)).astype(np.float32)
    tex, keep = make_texture([level])
    for y in range(4):
        for x in range(8):
            got = sample(tex, np.float32((x + 0.5) / 8), np.float32((y + 0.5) / 4))
            assert np.array_equal(got, level[y, x]), (x, y)


def test_two_by_two_hand_evaluation():
    level = np.array([[[0, 0, 0, 1], [1, 0, 0, 1]], [[0, 1, 0, 1], [1, 1, 1, 1]]], np.float32)
    tex, keep = make_texture([level], address_u=A["CLAMP_TO_EDGE"], address_v=A["CLAMP_TO_EDGE"])
    # u = 0.625: x = 0.75 texels from texel 0 -> t = 0.75; v = 0.375: y = 0.25 -> t = 0.25
    got = sample(tex, np.float32(0.625), np.float32(0.375))
    # rows: y0 row lerp (0,0,0,1)->(1,0,0,1) at .75 = (.75,0,0,1); y1 row (0,1,0,1)->(1,1,1,1) = (.75,1,.75,1)
    # columns at .25: (.75, .25, .1875, 1)
    assert np.array_equal(got, np.array([0.75, 0.25, 0.1875, 1.0], np.float32))


def _wrap_expected(i, n, mode):
    if mode == A["REPEAT"]:
        return ((i % n) + n) % n
    if mode == A["MIRRORED_REPEAT"]:
        # mirror((i) mod 2n): tiles alternate direction
        m = ((i % (2 * n)) + 2 * n) % (2 * n)
        return m if m < n else 2 * n - 1 - m
    if mode == A["CLAMP_TO_EDGE"]:
        return min(max(i, 0), n - 1)
    if mode == A["MIRROR_CLAMP_TO_EDGE"]:
        j = -1 - i if i < 0 else i
        return min(j, n - 1)
    return i if 0 <= i < n else None     # CLAMP_TO_BORDER


@pytest.mark.parametrize("mode", ["REPEAT", "MIRRORED_REPEAT", "CLAMP_TO_EDGE", "CLAMP_TO_BORDER",
                                  "MIRROR_CLAMP_TO_EDGE"])
@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_wrap_tables_nearest(mode, n):
    level = np.zeros((1, n, 4), np.float32)
    level[0, :, 0] = np.arange(n) + 10.0
    level[0, :, 3] = 1.0
    tex, keep = make_texture([level], mag=FL["NEAREST"], min=FL["NEAREST"], address_u=A[mode],
                             address_v=A["CLAMP_TO_EDGE"], border=3)
    for i in range(-3 * n, 3 * n):
        u = np.float32((i + 0.5) / n)
        got = sample(tex, u, np.float32(0.5))
        j = _wrap_expected(i, n, A[mode])
        if j is None:
            assert np.array_equal(got, np.array([0, 0, 0, 1], np.float32)), (i, got)   # opaque black border
        else:
            assert got[0] == j + 10.0, (mode, n, i, got[0], j)


def test_implicit_lod_ignores_other_levels():
    rng = np.random.default_rng(2)
    base = rng.uniform(0, 1, size=(8, 8, 4)).astype(np.float32)
    a, ka = make_texture(mips_of(base))
    lv = mips_of(base)
    for l in lv[1:]:
        l[...] = 1e6
    b, kb = make_texture(lv)
    c, kc = make_texture([base])
    for _ in range(200):
        u, v = np.float32(rng.uniform(-2, 2)), np.float32(rng.uniform(-2, 2))
        ra, rb, rc = sample(a, u, v), sample(b, u, v), sample(c, u, v)
        assert np.array_equal(ra, rb) and np.array_equal(ra, rc)


def test_bias_selects_levels():
    base = np.full((4, 4, 4), 1.0, np.float32)
    lv = [base, np.full((2, 2, 4), 3.0, np.float32), np.full((1, 1, 4), 7.0, np.float32)]
    # MIPMAP_NEAREST: level = ceil(lambda + 0.5) - 1
    for bias, want in [(0.0, 1.0), (0.4, 1.0), (0.6, 3.0), (1.0, 3.0), (1.6, 7.0), (5.0, 7.0)]:
        t, k = make_texture(lv, mipmap=0, bias=bias)
        assert sample(t, np.float32(0.3), np.float32(0.6))[0] == want, bias
    # MIPMAP_LINEAR: fp64 lerp between floor(lambda) and the next level
    for bias, want in [(0.5, 2.0), (0.25, 1.5), (1.5, 5.0), (2.0, 7.0)]:
        t, k = make_texture(lv, mipmap=1, bias=bias)
        assert sample(t, np.float32(0.3), np.float32(0.6))[0] == want, bias
    # max_lod clamps the bias
    t, k = make_texture(lv, mipmap=1, bias=2.0, max_lod=1.0)
    assert sample(t, np.float32(0.3), np.float32(0.6))[0] == 3.0


def test_unorm_texture_matches_float_texture():
    rng = np.random.default_rng(3)
    codes = rng.integers(0, 256, size=(8, 8, 4), dtype=np.uint8)
    tu, ku = make_texture([codes], fmt=F["R8G8B8A8_UNORM"])
    tf, kf = make_texture([(codes.astype(np.float32) / np.float32(255.0)).astype(np.float32)])
    for _ in range(200):
        u, v = np.float32(rng.uniform(-1, 2)), np.float32(rng.uniform(-1, 2))
        assert np.array_equal(sample(tu, u, v), sample(tf, u, v))
