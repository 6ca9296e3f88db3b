"""Transfer, clear and blit commands behind the Vulkan ICD (GPU only).

tests/icd/icd_transfer.cpp records, in one command buffer:
   1  vkCmdCopyBuffer      staging (host) -> device buffer, 16384 bytes
   2  vkCmdCopyBuffer      device buffer -> out (host): (100->0, 1000), (5000->2000, 333), (16000->8000, 384)
   3  vkCmdClearColorImage image A (R8G8B8A8_UNORM 40x30, 3 levels, device) to (0.25, 0.5, 0.75, 1), every level
   4  vkCmdCopyBufferToImage device buffer -> A: level 0 whole from 0; level 1 rows of 12 at (3, 2), 9 rows,
                           bufferRowLength 25, from 7000; level 2 whole from 9000
   5  vkCmdClearColorImage image B (40x30) to (1, 0, 0, 1)
   6  vkCmdCopyImage       A level 0 (5, 4) -> B (20, 10), 15 x 12
   7  vkCmdClearColorImage image C (64x48) to 0
   8  vkCmdBlitImage       A level 0 (0,0)-(40,30) -> C (0,0)-(64,48); A level 1 (0,0)-(20,15) -> C (60,40)-(30,10); LINEAR
   9  vkCmdClearDepthStencilImage image D (D32_SFLOAT 16x16) to depth 0.375
  10  vkCmdFillBuffer      fill (host, 4096): (8, 1000, 0xCAFEBABE), (2048, VK_WHOLE_SIZE, 0x01020304)
  11  vkCmdUpdateBuffer    fill at 1500: bytes 0..63
  12  vkCmdCopyImageToBuffer A's three levels -> readA at 0, 4800, 6000
  13  vkCmdCopyImageToBuffer B, C, D -> readB, readC, readD
  14  vkCmdCopyImageToBuffer A level 0, (2, 3), 10 x 5 -> readE at 100 with bufferRowLength 16
and every result must equal the oracle's bytes (the row restatement of
oracle/transfer_ref.py, the C oracle's clears and blits)."""
from __future__ import annotations

import ctypes as C
import struct
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib
from cpvulkan_amd import abi

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import transfer_ref as T  # noqa: E402

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
APP = ROOT / "tests" / "icd" / "build" / "icd_transfer"
ICD = ROOT / "cpvulkan_amd" / "lib" / "libcpvulkan_amd_icd.so"
RGBA8 = abi.FORMAT["R8G8B8A8_UNORM"]


def f32(*v):
    return [struct.unpack("<I", struct.pack("<f", x))[0] for x in v]


def expected(src: np.ndarray) -> dict:
    e = {}
    out = np.zeros(16384, np.uint8)
    T.copy_buffer(src, out, [(100, 0, 1000), (5000, 2000, 333), (16000, 8000, 384)])
    e["outbuf"] = out
    # image A: three levels 40x30, 20x15, 10x7 packed one after another
    dims = [(40, 30), (20, 15), (10, 7)]
    A = [np.zeros(w * h * 4, np.uint8) for w, h in dims]
    for lv, (w, h) in zip(A, dims):
        oracle_lib.clear_rects(lv, RGBA8, w, h, f32(0.25, 0.5, 0.75, 1.0), [abi.Rect(0, 0, w, h)])
    T.buffer_to_image(src, A[0], 40, 4, (0, 0, (0, 0), (40, 30)))
    T.buffer_to_image(src, A[1], 20, 4, (7000, 25, (3, 2), (12, 9)))
    T.buffer_to_image(src, A[2], 10, 4, (9000, 0, (0, 0), (10, 7)))
    e["readA"] = np.concatenate(A)
    B = np.zeros(40 * 30 * 4, np.uint8)
    oracle_lib.clear_rects(B, RGBA8, 40, 30, f32(1.0, 0.0, 0.0, 1.0), [abi.Rect(0, 0, 40, 30)])
    T.copy_image(A[0], 40, B, 40, 4, (5, 4), (20, 10), (15, 12))
    e["readB"] = B
    Cimg = np.zeros(64 * 48 * 4, np.uint8)
    r0 = abi.BlitRegion((C.c_int32 * 4)(0, 0, 40, 30), (C.c_int32 * 4)(0, 0, 64, 48), 1, 0)
    oracle_lib.blit_image(A[0].copy(), RGBA8, 40, 30, Cimg, RGBA8, 64, 48, [r0], 1)
    r1 = abi.BlitRegion((C.c_int32 * 4)(0, 0, 20, 15), (C.c_int32 * 4)(60, 40, 30, 10), 1, 0)
    oracle_lib.blit_image(A[1].copy(), RGBA8, 20, 15, Cimg, RGBA8, 64, 48, [r1], 1)
    e["readC"] = Cimg
    D = np.zeros(16 * 16 * 4, np.uint8)
    oracle_lib.clear_rects(D, abi.FORMAT["D32_SFLOAT"], 16, 16, f32(0.375) + [0, 0, 0], [abi.Rect(0, 0, 16, 16)])
    e["readD"] = D
    E = np.zeros(1024, np.uint8)
    T.image_to_buffer(A[0], 40, 4, E, (100, 16, (2, 3), (10, 5)))
    e["readE"] = E
    fill = np.zeros(4096, np.uint8)
    oracle_lib.fill_buffer(fill, 1000, 0xCAFEBABE, 8)
    oracle_lib.fill_buffer(fill, 2048, 0x01020304, 2048)
    fill[1500:1564] = np.arange(64, dtype=np.uint8)
    e["fill"] = fill
    return e


def test_transfer_commands_through_the_icd(tmp_path):
    src = np.random.Generator(np.random.PCG64(11)).integers(0, 256, size=16384, dtype=np.uint8)
    src.tofile(tmp_path / "src.bin")
    print(f"[icd] running {APP.name}", flush=True)
    res = subprocess.run([str(APP), str(ICD), str(tmp_path)], capture_output=True, text=True, timeout=240)
    print(res.stdout + res.stderr, flush=True)
    assert res.returncode == 0
    want = expected(src)
    bad = []
    for name, ref in want.items():
        got = np.fromfile(tmp_path / f"{name}.bin", np.uint8)
        if not np.array_equal(got, ref):
            bad.append(f"{name}: {np.count_nonzero(got != ref)} of {ref.size} bytes differ (first at {np.argmax(got != ref)})")
    assert not bad, "\n".join(bad)
