"""The Vulkan ICD's loader interface on the CPU: the three exported entry
points, interface-version negotiation (clamped to 5 as the reference does,
CPVulkan/CPVulkan.cpp:95-104), name lookup, and the instance-level queries
that need no device."""
import ctypes as C
import json
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
ICD = ROOT / "cpvulkan_amd" / "lib" / "libcpvulkan_amd_icd.so"


@pytest.fixture(scope="module")
def icd():
    if not ICD.exists():
        pytest.fail(f"{ICD} is not built; run python -m cpvulkan_amd.build icd")
    lib = C.CDLL(str(ICD))
    lib.vk_icdGetInstanceProcAddr.restype = C.c_void_p
    lib.vk_icdGetInstanceProcAddr.argtypes = [C.c_void_p, C.c_char_p]
    lib.vk_icdGetPhysicalDeviceProcAddr.restype = C.c_void_p
    lib.vk_icdGetPhysicalDeviceProcAddr.argtypes = [C.c_void_p, C.c_char_p]
    return lib


def test_exports_only_the_icd_interface():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", str(ICD)], capture_output=True, text=True, check=True).stdout
    names = sorted(l.split()[-1] for l in out.splitlines() if " T " in l)
    assert names == ["vk_icdGetInstanceProcAddr", "vk_icdGetPhysicalDeviceProcAddr", "vk_icdNegotiateLoaderICDInterfaceVersion"]


def test_manifest_points_at_the_library():
    m = json.loads((ROOT / "cpvulkan_amd" / "cpvulkan_amd_icd.json").read_text())
    assert m["file_format_version"] == "1.0.0"
    assert (ROOT / "cpvulkan_amd" / m["ICD"]["library_path"]).resolve() == ICD.resolve()


@pytest.mark.parametrize("asked,granted", [(6, 5), (5, 5), (3, 3)])
def test_negotiation(icd, asked, granted):
    v = C.c_uint32(asked)
    assert icd.vk_icdNegotiateLoaderICDInterfaceVersion(C.byref(v)) == 0
    assert v.value == granted


def test_proc_addr_lookup(icd):
    assert icd.vk_icdGetInstanceProcAddr(None, b"vkCreateInstance")
    assert icd.vk_icdGetInstanceProcAddr(None, b"vkCmdDraw") is None          # device-level names need an instance
    assert icd.vk_icdGetInstanceProcAddr(None, b"vkNoSuchFunction") is None
    assert icd.vk_icdGetPhysicalDeviceProcAddr(None, b"vkCmdDraw") is None


def test_instance_and_physical_device(icd):
    create = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p))(
        icd.vk_icdGetInstanceProcAddr(None, b"vkCreateInstance"))
    inst = C.c_void_p()
    info = (C.c_uint8 * 64)()
    C.cast(info, C.POINTER(C.c_int32))[0] = 1        # VK_STRUCTURE_TYPE_INSTANCE_CREATE_INFO, everything else zero
    assert create(C.addressof(info), None, C.byref(inst)) == 0
    # dispatchable handles start with the loader magic
    assert C.cast(inst, C.POINTER(C.c_uint64))[0] == 0x01CDC0DE
    for name in (b"vkCmdDraw", b"vkCmdDrawIndexed", b"vkQueueSubmit", b"vkCreateGraphicsPipelines", b"vkCmdDrawIndirect",
                 b"vkCmdDrawIndexedIndirect", b"vkCmdCopyBuffer", b"vkCmdCopyImage", b"vkCmdCopyBufferToImage",
                 b"vkCmdCopyImageToBuffer", b"vkCmdFillBuffer", b"vkCmdUpdateBuffer", b"vkCmdClearColorImage",
                 b"vkCmdClearDepthStencilImage", b"vkCmdClearAttachments", b"vkCmdBlitImage", b"vkCmdSetDepthBounds",
                 b"vkCmdSetLineWidth"):
        assert icd.vk_icdGetInstanceProcAddr(inst, name), name
    enum = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_void_p))(
        icd.vk_icdGetInstanceProcAddr(inst, b"vkEnumeratePhysicalDevices"))
    n = C.c_uint32(0)
    assert enum(inst, C.byref(n), None) == 0 and n.value == 1
    phys = C.c_void_p()
    assert enum(inst, C.byref(n), C.byref(phys)) == 0
    props = (C.c_uint8 * 824)()
    getp = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)(icd.vk_icdGetInstanceProcAddr(inst, b"vkGetPhysicalDeviceProperties"))
    getp(phys, C.addressof(props))
    api = C.cast(props, C.POINTER(C.c_uint32))[0]
    assert api == (1 << 22) | (1 << 12)
    name = bytes(props[20:276]).split(b"\0")[0].decode()
    assert "MI355X" in name
    destroy = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)(icd.vk_icdGetInstanceProcAddr(inst, b"vkDestroyInstance"))
    destroy(inst, None)
