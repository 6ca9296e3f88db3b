// A Vulkan application for the ICD transfer test: one command buffer of
// copies, clears, blits, fills and updates between host-visible and
// device-local buffers and images, replayed by vkQueueSubmit; every result
// is read back through host-visible buffers and written to <dir>.
//
//   icd_transfer <icd.so> <dir>
//
// <dir>/src.bin holds 16384 bytes; the program writes outbuf.bin, readA.bin
// (image A's three levels), readB.bin, readC.bin, readD.bin, readE.bin and
// fill.bin.  The command sequence is listed in tests/test_gpu_icd_transfer.py,
// which rebuilds every result with the oracle.
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "cpv_vk_enums.h"
#include "cpv_vulkan.h"

#define CHECK(x)                                                                  \
    do {                                                                          \
        VkResult r_ = (x);                                                        \
        if (r_ != VK_SUCCESS) {                                                   \
            std::fprintf(stderr, "%s failed: %d (line %d)\n", #x, r_, __LINE__); \
            std::exit(2);                                                         \
        }                                                                         \
    } while (0)

static PFN_vkGetInstanceProcAddr gipa;
static VkInstance g_inst;

template <class F> F load(const char* name)
{
    F f = reinterpret_cast<F>(gipa(g_inst, name));
    if (!f) {
        std::fprintf(stderr, "missing entry point %s\n", name);
        std::exit(2);
    }
    return f;
}

#define FN(name, ...) auto vk##name = load<VkResult (*)(__VA_ARGS__)>("vk" #name)
#define VFN(name, ...) auto vk##name = load<void (*)(__VA_ARGS__)>("vk" #name)

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: icd_transfer <icd.so> <dir>\n");
        return 2;
    }
    const std::string dir = argv[2];
    void* lib = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if (!lib) {
        std::fprintf(stderr, "dlopen: %s\n", dlerror());
        return 2;
    }
    auto negotiate = reinterpret_cast<VkResult (*)(uint32_t*)>(dlsym(lib, "vk_icdNegotiateLoaderICDInterfaceVersion"));
    gipa = reinterpret_cast<PFN_vkGetInstanceProcAddr>(dlsym(lib, "vk_icdGetInstanceProcAddr"));
    if (!negotiate || !gipa) return 2;
    uint32_t ver = 5;
    CHECK(negotiate(&ver));

    auto vkCreateInstance = reinterpret_cast<VkResult (*)(const VkInstanceCreateInfo*, const VkAllocationCallbacks*, VkInstance*)>(
        gipa(nullptr, "vkCreateInstance"));
    VkApplicationInfo app{VK_STRUCTURE_TYPE_APPLICATION_INFO, nullptr, "icd_transfer", 1, "none", 1, VK_API_VERSION_1_1};
    VkInstanceCreateInfo ici{VK_STRUCTURE_TYPE_INSTANCE_CREATE_INFO, nullptr, 0, &app, 0, nullptr, 0, nullptr};
    CHECK(vkCreateInstance(&ici, nullptr, &g_inst));

    FN(EnumeratePhysicalDevices, VkInstance, uint32_t*, VkPhysicalDevice*);
    VFN(GetPhysicalDeviceMemoryProperties, VkPhysicalDevice, VkPhysicalDeviceMemoryProperties*);
    FN(CreateDevice, VkPhysicalDevice, const VkDeviceCreateInfo*, const VkAllocationCallbacks*, VkDevice*);
    VFN(GetDeviceQueue, VkDevice, uint32_t, uint32_t, VkQueue*);
    FN(AllocateMemory, VkDevice, const VkMemoryAllocateInfo*, const VkAllocationCallbacks*, VkDeviceMemory*);
    FN(MapMemory, VkDevice, VkDeviceMemory, VkDeviceSize, VkDeviceSize, VkFlags, void**);
    FN(CreateBuffer, VkDevice, const VkBufferCreateInfo*, const VkAllocationCallbacks*, VkBuffer*);
    VFN(GetBufferMemoryRequirements, VkDevice, VkBuffer, VkMemoryRequirements*);
    FN(BindBufferMemory, VkDevice, VkBuffer, VkDeviceMemory, VkDeviceSize);
    FN(CreateImage, VkDevice, const VkImageCreateInfo*, const VkAllocationCallbacks*, VkImage*);
    VFN(GetImageMemoryRequirements, VkDevice, VkImage, VkMemoryRequirements*);
    FN(BindImageMemory, VkDevice, VkImage, VkDeviceMemory, VkDeviceSize);
    FN(CreateCommandPool, VkDevice, const VkCommandPoolCreateInfo*, const VkAllocationCallbacks*, VkCommandPool*);
    FN(AllocateCommandBuffers, VkDevice, const VkCommandBufferAllocateInfo*, VkCommandBuffer*);
    FN(BeginCommandBuffer, VkCommandBuffer, const VkCommandBufferBeginInfo*);
    FN(EndCommandBuffer, VkCommandBuffer);
    VFN(CmdCopyBuffer, VkCommandBuffer, VkBuffer, VkBuffer, uint32_t, const VkBufferCopy*);
    VFN(CmdCopyImage, VkCommandBuffer, VkImage, int32_t, VkImage, int32_t, uint32_t, const VkImageCopy*);
    VFN(CmdCopyBufferToImage, VkCommandBuffer, VkBuffer, VkImage, int32_t, uint32_t, const VkBufferImageCopy*);
    VFN(CmdCopyImageToBuffer, VkCommandBuffer, VkImage, int32_t, VkBuffer, uint32_t, const VkBufferImageCopy*);
    VFN(CmdFillBuffer, VkCommandBuffer, VkBuffer, VkDeviceSize, VkDeviceSize, uint32_t);
    VFN(CmdUpdateBuffer, VkCommandBuffer, VkBuffer, VkDeviceSize, VkDeviceSize, const void*);
    VFN(CmdClearColorImage, VkCommandBuffer, VkImage, int32_t, const VkClearColorValue*, uint32_t, const VkImageSubresourceRange*);
    VFN(CmdClearDepthStencilImage, VkCommandBuffer, VkImage, int32_t, const VkClearDepthStencilValue*, uint32_t,
        const VkImageSubresourceRange*);
    VFN(CmdBlitImage, VkCommandBuffer, VkImage, int32_t, VkImage, int32_t, uint32_t, const VkImageBlit*, int32_t);
    FN(QueueSubmit, VkQueue, uint32_t, const VkSubmitInfo*, VkFence);
    FN(CreateFence, VkDevice, const VkFenceCreateInfo*, const VkAllocationCallbacks*, VkFence*);
    FN(WaitForFences, VkDevice, uint32_t, const VkFence*, VkBool32, uint64_t);
    VFN(DestroyDevice, VkDevice, const VkAllocationCallbacks*);
    VFN(DestroyInstance, VkInstance, const VkAllocationCallbacks*);

    uint32_t n = 1;
    VkPhysicalDevice phys;
    CHECK(vkEnumeratePhysicalDevices(g_inst, &n, &phys));
    VkPhysicalDeviceMemoryProperties mp;
    vkGetPhysicalDeviceMemoryProperties(phys, &mp);
    uint32_t host_type = ~0u, device_type = ~0u;
    for (uint32_t i = 0; i < mp.memoryTypeCount; i++) {
        if ((mp.memoryTypes[i].propertyFlags & VK_MEMORY_PROPERTY_HOST_VISIBLE_BIT) && host_type == ~0u) host_type = i;
        if (!(mp.memoryTypes[i].propertyFlags & VK_MEMORY_PROPERTY_HOST_VISIBLE_BIT) && device_type == ~0u) device_type = i;
    }
    if (host_type == ~0u || device_type == ~0u) return 4;

    const float prio = 1.0f;
    VkDeviceQueueCreateInfo qci{VK_STRUCTURE_TYPE_DEVICE_QUEUE_CREATE_INFO, nullptr, 0, 0, 1, &prio};
    VkDeviceCreateInfo dci{VK_STRUCTURE_TYPE_DEVICE_CREATE_INFO, nullptr, 0, 1, &qci, 0, nullptr, 0, nullptr, nullptr};
    VkDevice dev;
    CHECK(vkCreateDevice(phys, &dci, nullptr, &dev));
    VkQueue queue;
    vkGetDeviceQueue(dev, 0, 0, &queue);

    // a buffer in host-visible (mapped, returned in *map) or device-local memory
    auto buffer = [&](VkDeviceSize size, bool host, VkBuffer* buf, void** map) {
        VkBufferCreateInfo bi{VK_STRUCTURE_TYPE_BUFFER_CREATE_INFO, nullptr, 0, size, 0x3, 0, 0, nullptr};
        CHECK(vkCreateBuffer(dev, &bi, nullptr, buf));
        VkMemoryRequirements req;
        vkGetBufferMemoryRequirements(dev, *buf, &req);
        VkMemoryAllocateInfo ai{VK_STRUCTURE_TYPE_MEMORY_ALLOCATE_INFO, nullptr, req.size, host ? host_type : device_type};
        VkDeviceMemory mem;
        CHECK(vkAllocateMemory(dev, &ai, nullptr, &mem));
        if (host) {
            CHECK(vkMapMemory(dev, mem, 0, req.size, 0, map));
            std::memset(*map, 0, size);
        }
        CHECK(vkBindBufferMemory(dev, *buf, mem, 0));
    };
    auto image = [&](int32_t format, uint32_t w, uint32_t h, uint32_t levels, VkImage* img) {
        VkImageCreateInfo ii{VK_STRUCTURE_TYPE_IMAGE_CREATE_INFO, nullptr, 0, VK_IMAGE_TYPE_2D, format, {w, h, 1}, levels, 1, 1,
                             VK_IMAGE_TILING_LINEAR, 0x1 | 0x2, 0, 0, nullptr, 0};
        CHECK(vkCreateImage(dev, &ii, nullptr, img));
        VkMemoryRequirements req;
        vkGetImageMemoryRequirements(dev, *img, &req);
        VkMemoryAllocateInfo ai{VK_STRUCTURE_TYPE_MEMORY_ALLOCATE_INFO, nullptr, req.size, device_type};
        VkDeviceMemory mem;
        CHECK(vkAllocateMemory(dev, &ai, nullptr, &mem));
        CHECK(vkBindImageMemory(dev, *img, mem, 0));
    };

    std::ifstream sf(dir + "/src.bin", std::ios::binary);
    std::vector<char> src((std::istreambuf_iterator<char>(sf)), std::istreambuf_iterator<char>());
    if (src.size() != 16384) return 5;

    VkBuffer stage, devbuf, outbuf, readA, readB, readC, readD, readE, fill;
    void *stage_map, *out_map, *a_map, *b_map, *c_map, *d_map, *e_map, *fill_map, *none = nullptr;
    buffer(16384, true, &stage, &stage_map);
    std::memcpy(stage_map, src.data(), src.size());
    buffer(16384, false, &devbuf, &none);
    buffer(16384, true, &outbuf, &out_map);
    buffer(6280, true, &readA, &a_map);
    buffer(4800, true, &readB, &b_map);
    buffer(64 * 48 * 4, true, &readC, &c_map);
    buffer(16 * 16 * 4, true, &readD, &d_map);
    buffer(1024, true, &readE, &e_map);
    buffer(4096, true, &fill, &fill_map);
    VkImage imgA, imgB, imgC, imgD;
    image(CPV_FORMAT_R8G8B8A8_UNORM, 40, 30, 3, &imgA);
    image(CPV_FORMAT_R8G8B8A8_UNORM, 40, 30, 1, &imgB);
    image(CPV_FORMAT_R8G8B8A8_UNORM, 64, 48, 1, &imgC);
    image(CPV_FORMAT_D32_SFLOAT, 16, 16, 1, &imgD);

    VkCommandPoolCreateInfo cpi{VK_STRUCTURE_TYPE_COMMAND_POOL_CREATE_INFO, nullptr, 0, 0};
    VkCommandPool cpool;
    CHECK(vkCreateCommandPool(dev, &cpi, nullptr, &cpool));
    VkCommandBufferAllocateInfo cai{VK_STRUCTURE_TYPE_COMMAND_BUFFER_ALLOCATE_INFO, nullptr, cpool, VK_COMMAND_BUFFER_LEVEL_PRIMARY, 1};
    VkCommandBuffer cmd;
    CHECK(vkAllocateCommandBuffers(dev, &cai, &cmd));
    VkCommandBufferBeginInfo cbi{VK_STRUCTURE_TYPE_COMMAND_BUFFER_BEGIN_INFO, nullptr, 0, nullptr};
    CHECK(vkBeginCommandBuffer(cmd, &cbi));

    const VkImageSubresourceRange all_levels{VK_IMAGE_ASPECT_COLOR_BIT, 0, VK_REMAINING_MIP_LEVELS, 0, 1};
    const VkImageSubresourceRange level0{VK_IMAGE_ASPECT_COLOR_BIT, 0, 1, 0, 1};
    const VkImageSubresourceLayers a0{VK_IMAGE_ASPECT_COLOR_BIT, 0, 0, 1}, a1{VK_IMAGE_ASPECT_COLOR_BIT, 1, 0, 1},
        a2{VK_IMAGE_ASPECT_COLOR_BIT, 2, 0, 1};
    // 1-2: staging -> device buffer -> three regions of a host buffer
    const VkBufferCopy whole{0, 0, 16384};
    vkCmdCopyBuffer(cmd, stage, devbuf, 1, &whole);
    const VkBufferCopy parts[3] = {{100, 0, 1000}, {5000, 2000, 333}, {16000, 8000, 384}};
    vkCmdCopyBuffer(cmd, devbuf, outbuf, 3, parts);
    // 3-4: image A cleared, then filled from the device buffer
    VkClearColorValue ca;
    ca.float32[0] = 0.25f; ca.float32[1] = 0.5f; ca.float32[2] = 0.75f; ca.float32[3] = 1.0f;
    vkCmdClearColorImage(cmd, imgA, 1, &ca, 1, &all_levels);
    const VkBufferImageCopy up[3] = {{0, 0, 0, a0, {0, 0, 0}, {40, 30, 1}},
                                     {7000, 25, 0, a1, {3, 2, 0}, {12, 9, 1}},
                                     {9000, 0, 0, a2, {0, 0, 0}, {10, 7, 1}}};
    vkCmdCopyBufferToImage(cmd, devbuf, imgA, 1, 3, up);
    // 5-6: image B cleared to red, a rectangle of A's level 0 copied in
    VkClearColorValue cb;
    cb.float32[0] = 1.0f; cb.float32[1] = 0.0f; cb.float32[2] = 0.0f; cb.float32[3] = 1.0f;
    vkCmdClearColorImage(cmd, imgB, 1, &cb, 1, &level0);
    const VkImageCopy ic{a0, {5, 4, 0}, a0, {20, 10, 0}, {15, 12, 1}};
    vkCmdCopyImage(cmd, imgA, 1, imgB, 1, 1, &ic);
    // 7-8: image C cleared, then blitted from A's levels 0 and 1 (scaled, mirrored)
    VkClearColorValue cc;
    std::memset(&cc, 0, sizeof(cc));
    vkCmdClearColorImage(cmd, imgC, 1, &cc, 1, &level0);
    const VkImageBlit bl[2] = {{a0, {{0, 0, 0}, {40, 30, 1}}, a0, {{0, 0, 0}, {64, 48, 1}}},
                               {a1, {{0, 0, 0}, {20, 15, 1}}, a0, {{60, 40, 0}, {30, 10, 1}}}};
    vkCmdBlitImage(cmd, imgA, 1, imgC, 1, 2, bl, VK_FILTER_LINEAR);
    // 9: depth image cleared
    const VkClearDepthStencilValue dv{0.375f, 0};
    const VkImageSubresourceRange depth0{VK_IMAGE_ASPECT_DEPTH_BIT, 0, 1, 0, 1};
    vkCmdClearDepthStencilImage(cmd, imgD, 1, &dv, 1, &depth0);
    // 10-11: fills and an update
    vkCmdFillBuffer(cmd, fill, 8, 1000, 0xCAFEBABEu);
    vkCmdFillBuffer(cmd, fill, 2048, VK_WHOLE_SIZE, 0x01020304u);
    uint8_t upd[64];
    for (int i = 0; i < 64; i++) upd[i] = (uint8_t)i;
    vkCmdUpdateBuffer(cmd, fill, 1500, 64, upd);
    // 12-14: read back
    const VkBufferImageCopy downA[3] = {{0, 0, 0, a0, {0, 0, 0}, {40, 30, 1}},
                                        {4800, 0, 0, a1, {0, 0, 0}, {20, 15, 1}},
                                        {6000, 0, 0, a2, {0, 0, 0}, {10, 7, 1}}};
    vkCmdCopyImageToBuffer(cmd, imgA, 1, readA, 3, downA);
    const VkBufferImageCopy downB{0, 0, 0, a0, {0, 0, 0}, {40, 30, 1}};
    vkCmdCopyImageToBuffer(cmd, imgB, 1, readB, 1, &downB);
    const VkBufferImageCopy downC{0, 0, 0, a0, {0, 0, 0}, {64, 48, 1}};
    vkCmdCopyImageToBuffer(cmd, imgC, 1, readC, 1, &downC);
    const VkBufferImageCopy downD{0, 0, 0, {VK_IMAGE_ASPECT_DEPTH_BIT, 0, 0, 1}, {0, 0, 0}, {16, 16, 1}};
    vkCmdCopyImageToBuffer(cmd, imgD, 1, readD, 1, &downD);
    const VkBufferImageCopy downE{100, 16, 0, a0, {2, 3, 0}, {10, 5, 1}};
    vkCmdCopyImageToBuffer(cmd, imgA, 1, readE, 1, &downE);
    CHECK(vkEndCommandBuffer(cmd));

    VkFenceCreateInfo fci{VK_STRUCTURE_TYPE_FENCE_CREATE_INFO, nullptr, 0};
    VkFence fence;
    CHECK(vkCreateFence(dev, &fci, nullptr, &fence));
    VkSubmitInfo si{VK_STRUCTURE_TYPE_SUBMIT_INFO, nullptr, 0, nullptr, nullptr, 1, &cmd, 0, nullptr};
    CHECK(vkQueueSubmit(queue, 1, &si, fence));
    CHECK(vkWaitForFences(dev, 1, &fence, VK_TRUE, ~0ull));

    auto out = [&](const char* name, const void* p, size_t n) { std::ofstream(dir + "/" + name, std::ios::binary).write((const char*)p, (std::streamsize)n); };
    out("outbuf.bin", out_map, 16384);
    out("readA.bin", a_map, 6280);
    out("readB.bin", b_map, 4800);
    out("readC.bin", c_map, 64 * 48 * 4);
    out("readD.bin", d_map, 16 * 16 * 4);
    out("readE.bin", e_map, 1024);
    out("fill.bin", fill_map, 4096);
    vkDestroyDevice(dev, nullptr);
    vkDestroyInstance(g_inst, nullptr);
    std::printf("transfers done\n");
    return 0;
}
