"""Every operation the SPIR-V translator emits, on the GPU, against the CPU
oracle of its semantics (oracle/spirv_alu.py).

The conformance pair (cpvulkan_amd.spirv.vs_flat_inputs + fs_alu_ops) is
translated and compiled with hiprtc, and drawn over a 128x64
R32G32B32A32_SFLOAT target as one small triangle per pixel centre: each
triangle carries, as flat varyings, the inputs a, b, c, d of one case and
the case number, so every fragment evaluates one operation on inputs the
test chose and stores its exact result bits.  Results must match the
oracle bit for bit, except for the transcendental functions, which must lie
within oracle.spirv_alu.TOLERANCE_ULPS of the correctly rounded value."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

import cpvulkan_amd
from cpvulkan_amd import abi
from cpvulkan_amd import spirv as S
from cpvulkan_amd.workloads import DeviceMemory, DrawData, Scene, materialize

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import spirv_alu as O  # noqa: E402

pytestmark = pytest.mark.gpu

W, H = 128, 64
SFLOAT4 = abi.FORMAT["R32G32B32A32_SFLOAT"]


def _domain(name: str, rng, n: int):
    """Inputs (a, b, c, d), each (n, 4) float32, inside the case's domain."""
    def u(lo, hi):
        return rng.uniform(lo, hi, size=(n, 4)).astype(np.float32)

    def bits():
        return rng.integers(0, 2**32, size=(n, 4), dtype=np.uint64).astype(np.uint32).view(np.float32)

    def mixed():
        x = u(-10, 10)
        k = rng.random((n, 4))
        x[k < 0.15] = rng.integers(-8, 9, size=int((k < 0.15).sum())).astype(np.float32) + np.float32(0.5)
        x[(k >= 0.15) & (k < 0.25)] = rng.integers(-8, 9, size=int(((k >= 0.15) & (k < 0.25)).sum())).astype(np.float32)
        x[(k >= 0.25) & (k < 0.28)] = np.float32(-0.0)
        x[(k >= 0.28) & (k < 0.31)] = np.float32(0.0)
        x[(k >= 0.31) & (k < 0.35)] = u(-1e-3, 1e-3)[(k >= 0.31) & (k < 0.35)]
        return x

    def with_nan(x):
        x = x.copy()
        x[rng.random((n, 4)) < 0.15] = np.float32("nan")
        return x

    a, b, c, d = mixed(), mixed(), mixed(), mixed()
    if name in ("Asin", "Acos"):
        a = u(-1, 1)
    elif name == "Atanh":
        a = u(-0.999, 0.999)
    elif name == "Acosh":
        a = u(1, 100)
    elif name in ("Log", "Log2", "Sqrt", "InverseSqrt"):
        a = np.abs(a) + np.float32(1e-3)
    elif name in ("Exp", "Exp2", "Sinh", "Cosh"):
        a = u(-20, 20)
    elif name == "Pow":
        a, b = u(0.01, 10), u(-5, 5)
    elif name in ("NMin", "NMax", "NClamp", "FMin", "FMax", "FClamp", "FOrdCompare", "FUnordCompare", "IsNanAnyAll"):
        a, b = with_nan(a), with_nan(b)
        eq = rng.random((n, 4)) < 0.2
        b[eq] = a[eq]
        a[rng.random((n, 4)) < 0.05] = np.float32("inf")
    elif name in ("FDiv", "FRem", "FMod"):
        b[np.abs(b) < 1e-2] = np.float32(3.0)
    elif name == "Refract":
        c = u(0.5, 2.0)
    elif name == "Ldexp":
        b = u(-20, 20)
    elif name == "Convert":
        a = u(-1e6, 1e6)
        b = bits()
    elif name == "DynamicInsertExtract":
        b, c = u(-8, 8), u(-8, 8)
    elif name == "Pack":
        a = u(-1.5, 1.5)
        a[rng.random((n, 4)) < 0.2] = np.float32(0.5) / np.float32(255.0)
    elif name in ("Unpack4x8", "Unpack2x16"):
        a = bits()
        # keep the half-precision fields finite (exponent field below 31)
        ab = a.view(np.uint32)
        ab &= np.uint32(0xBBFFBBFF)
    elif name in ("SDiv", "SRem", "SMod", "UDiv", "UMod"):
        a = bits()
        bi = rng.integers(1, 2**31 - 1, size=(n, 4)).astype(np.int64)
        small = rng.random((n, 4)) < 0.5
        bi[small] = rng.integers(1, 100, size=int(small.sum()))
        bi[rng.random((n, 4)) < 0.5] *= -1
        b = (bi & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    elif name in ("ShiftLeftLogical", "ShiftRightLogical", "ShiftRightArithmetic"):
        a = bits()
        b = rng.integers(0, 32, size=(n, 4)).astype(np.uint32).view(np.float32)
    elif name in ("IAdd", "ISub", "IMul", "BitwiseAnd", "BitwiseOr", "BitwiseXor", "SNegate", "Not", "BitReverse",
                  "BitCount", "SAbs", "SSign", "FindILsb", "FindSMsb", "FindUMsb", "SMin", "SMax", "UMin", "UMax",
                  "SClamp", "UClamp", "BitFieldInsert", "BitFieldUExtract", "BitFieldSExtract", "IntCompare"):
        a, b, c = bits(), bits(), bits()
        # small magnitudes, zeros and equal pairs as well as random words
        sm = rng.random((n, 4)) < 0.3
        a.view(np.int32)[sm] = rng.integers(-40, 40, size=int(sm.sum()))
        a.view(np.uint32)[rng.random((n, 4)) < 0.05] = 0
        eq = rng.random((n, 4)) < 0.15
        b[eq] = a[eq]
    return a, b, c, d


def _scene(cases: np.ndarray, a, b, c, d, program) -> Scene:
    """One triangle around each pixel centre, flat inputs per triangle."""
    n = W * H
    px = np.arange(n)
    cx = ((px % W) + 0.5) / W * 2.0 - 1.0
    cy = ((px // W) + 0.5) / H * 2.0 - 1.0
    off = np.array([[-0.3, -0.3], [0.3, -0.3], [0.0, 0.3]])
    v = np.zeros((n, 3, 24), np.float32)
    v[:, :, 0] = cx[:, None] + off[None, :, 0] * (2.0 / W)
    v[:, :, 1] = cy[:, None] + off[None, :, 1] * (2.0 / H)
    v[:, :, 2] = 0.5
    v[:, :, 3] = 1.0
    for k, arr in enumerate((a, b, c, d)):
        v[:, :, 4 + 4 * k:8 + 4 * k] = arr[:, None, :]
    v[:, :, 20] = cases[:, None].astype(np.float32)
    vec4 = SFLOAT4
    pipeline = dict(
        program=program.handle.value, bindings=[(0, 96, 0)],
        attributes=[(loc, 0, vec4, 16 * loc) for loc in range(6)],
        cull=abi.CULL["NONE"], front_face=abi.FRONT_FACE["CCW"], depth_test=0, depth_write=0,
    )
    draw = DrawData(vertices={0: v.reshape(-1).view(np.uint8)}, count=n * 3,
                    viewport=(0.0, 0.0, float(W), float(H), 0.0, 1.0))
    return Scene("alu", W, H, SFLOAT4, 0, pipeline, [draw], clear_color=(-7.0, -7.0, -7.0, -7.0))


@pytest.fixture(scope="module")
def alu_program(gpu_ctx):
    print("[spirv-alu] creating the conformance program", flush=True)
    prog = cpvulkan_amd.Program(gpu_ctx, S.vs_flat_inputs(), S.fs_alu_ops())
    yield prog
    prog.close()


@pytest.mark.parametrize("seed", [1, 2])
def test_every_operation_matches_the_oracle(gpu_ctx, alu_program, seed):
    names = S.ALU_CASES
    rng = np.random.Generator(np.random.PCG64(1000 + seed))
    per = (W * H) // len(names)
    cases = np.minimum(np.arange(W * H) // per, len(names) - 1)
    a = np.zeros((W * H, 4), np.float32)
    b, c, d = a.copy(), a.copy(), a.copy()
    for k, name in enumerate(names):
        sel = cases == k
        a[sel], b[sel], c[sel], d[sel] = _domain(name, rng, int(sel.sum()))
    scene = _scene(cases, a, b, c, d, alu_program)
    print(f"[spirv-alu] seed {seed}: {len(names)} cases x {per} fragments", flush=True)
    mem = DeviceMemory(gpu_ctx)
    try:
        m = materialize(scene, mem)
        gpu_ctx.render_pass(m.pass_desc, m.pipelines, m.draws, m.draw_pipeline)
        gpu_ctx.synchronize()
        out = m.read_color().view(np.float32).reshape(W * H, 4)
    finally:
        mem.release()
    assert not np.any(np.all(out == np.float32(-7.0), axis=1)), "a fragment was not shaded"
    failures = []
    for k, name in enumerate(names):
        sel = cases == k
        want = O.evaluate(name, a[sel], b[sel], c[sel], d[sel])
        got = out[sel]
        tol = O.TOLERANCE_ULPS.get(name, 0)
        dist = O.ulp_distance(got, want)
        bad = dist > tol
        if bad.any():
            i = np.argwhere(bad)[0]
            r = i[0]
            failures.append(f"{name}: {int(bad.sum())} components beyond {tol} ulp; first a={a[sel][r]} b={b[sel][r]} "
                            f"c={c[sel][r]} gpu={got[r]} ({got[r].view(np.uint32)}) oracle={want[r]} ({want[r].view(np.uint32)})")
    assert not failures, "\n".join(failures)
