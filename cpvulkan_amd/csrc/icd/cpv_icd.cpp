// Vulkan ICD front end of the draw path: libcpvulkan_amd_icd.so.
//
// An application that points the loader at cpvulkan_amd_icd.json (or that
// dlopens this library and calls vk_icdGetInstanceProcAddr) gets a Vulkan 1.1
// device whose vkCmdDraw / vkCmdDrawIndexed, recorded into command buffers
// and replayed by vkQueueSubmit, run on the MI355X through the C-ABI of
// include/cpv.h.  The object model is the minimum the draw path needs; it
// mirrors the reference's (CPVulkan/*.cpp): commands are recorded into the
// command buffer and executed at submit (CommandBuffer.h:161,
// CommandBuffer.cpp:964-968).  A submit gathers the render passes of its
// command buffers and hands them to libcpv in one cpv_render_passes call
// (independent passes share launches); a command buffer recorded without
// ONE_TIME_SUBMIT whose commands are only render passes is recorded once as
// a libcpv batch (hipGraphs) and replayed by each submit; a VkFence is a HIP
// event recorded after the submit's work.
//
// Memory: memory type 0 is device-local HBM (cpv_alloc); memory type 1 is
// host-visible, host-coherent pinned memory (cpv_alloc_host) that kernels
// read and write directly, so vkMapMemory needs no copies.  Images are
// linear (the reference's layout: tightly packed rows, mip chain level after
// level, CPVulkanBase/Formats.cpp:455-587).
//
// Errors never abort: unsupported state returns VK_ERROR_FEATURE_NOT_PRESENT
// (where the reference calls TODO_ERROR), HIP failures VK_ERROR_DEVICE_LOST.
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/cpv.h"
#include "../../../include/cpv_vulkan.h"

namespace {

uint32_t format_bytes(int32_t f)
{
    switch (f) {
    case CPV_FORMAT_R8_UNORM: case CPV_FORMAT_R8_SNORM: case CPV_FORMAT_R8_USCALED: case CPV_FORMAT_R8_SSCALED:
    case CPV_FORMAT_R8_UINT: case CPV_FORMAT_R8_SINT: case CPV_FORMAT_S8_UINT: return 1;
    case CPV_FORMAT_R8G8_UNORM: case CPV_FORMAT_R8G8_SNORM: case CPV_FORMAT_R8G8_UINT: case CPV_FORMAT_R8G8_SINT:
    case CPV_FORMAT_R16_UNORM: case CPV_FORMAT_R16_SNORM: case CPV_FORMAT_R16_UINT: case CPV_FORMAT_R16_SINT:
    case CPV_FORMAT_D16_UNORM: return 2;
    case CPV_FORMAT_R8G8B8_UNORM: case CPV_FORMAT_R8G8B8_SNORM: case CPV_FORMAT_R8G8B8_UINT: case CPV_FORMAT_R8G8B8_SINT:
    case CPV_FORMAT_B8G8R8_UNORM: case CPV_FORMAT_B8G8R8_SNORM: case CPV_FORMAT_B8G8R8_UINT: case CPV_FORMAT_B8G8R8_SINT:
    case CPV_FORMAT_D16_UNORM_S8_UINT: return 3;
    case CPV_FORMAT_R8G8B8A8_UNORM: case CPV_FORMAT_R8G8B8A8_SNORM: case CPV_FORMAT_R8G8B8A8_USCALED:
    case CPV_FORMAT_R8G8B8A8_SSCALED: case CPV_FORMAT_R8G8B8A8_UINT: case CPV_FORMAT_R8G8B8A8_SINT:
    case CPV_FORMAT_B8G8R8A8_UNORM: case CPV_FORMAT_B8G8R8A8_SNORM: case CPV_FORMAT_B8G8R8A8_UINT: case CPV_FORMAT_B8G8R8A8_SINT:
    case CPV_FORMAT_R16G16_UNORM: case CPV_FORMAT_R16G16_UINT: case CPV_FORMAT_R16G16_SINT: case CPV_FORMAT_R32_UINT:
    case CPV_FORMAT_R32_SINT: case CPV_FORMAT_R32_SFLOAT: case CPV_FORMAT_X8_D24_UNORM_PACK32: case CPV_FORMAT_D32_SFLOAT:
    case CPV_FORMAT_D24_UNORM_S8_UINT: return 4;
    case CPV_FORMAT_R16G16B16A16_UNORM: case CPV_FORMAT_R16G16B16A16_SNORM: case CPV_FORMAT_R16G16B16A16_UINT:
    case CPV_FORMAT_R16G16B16A16_SINT: case CPV_FORMAT_R32G32_UINT: case CPV_FORMAT_R32G32_SINT: case CPV_FORMAT_R32G32_SFLOAT:
    case CPV_FORMAT_D32_SFLOAT_S8_UINT: return 8;
    case CPV_FORMAT_R32G32B32_UINT: case CPV_FORMAT_R32G32B32_SINT: case CPV_FORMAT_R32G32B32_SFLOAT: return 12;
    case CPV_FORMAT_R32G32B32A32_UINT: case CPV_FORMAT_R32G32B32A32_SINT: case CPV_FORMAT_R32G32B32A32_SFLOAT: return 16;
    default: return 0;
    }
}

bool is_depth_stencil(int32_t f) { return f >= CPV_FORMAT_D16_UNORM && f <= CPV_FORMAT_D32_SFLOAT_S8_UINT; }

// ------------------------------------------------------------------ objects

struct Dispatchable {
    uintptr_t loader_magic = ICD_LOADER_MAGIC;   // the loader replaces this with its dispatch table
};

struct Instance;
struct Device;

struct PhysicalDevice : Dispatchable {
    Instance* instance = nullptr;
};

struct Instance : Dispatchable {
    PhysicalDevice physical;
};

struct Queue : Dispatchable {
    Device* device = nullptr;
};

struct Device : Dispatchable {
    cpv_context* ctx = nullptr;
    Queue queue;
    std::mutex lock;
    uint64_t submits = 0;
};

struct Memory {
    uint64_t address = 0;
    VkDeviceSize size = 0;
    bool host = false;
};

struct Buffer {
    VkDeviceSize size = 0;
    Memory* memory = nullptr;
    VkDeviceSize offset = 0;
    uint64_t address() const { return memory ? memory->address + offset : 0; }
};

struct Image {
    int32_t format = 0;
    uint32_t width = 0, height = 0, levels = 1, layers = 1;
    VkDeviceSize size = 0;
    std::vector<VkDeviceSize> level_offset;
    Memory* memory = nullptr;
    VkDeviceSize offset = 0;
    uint64_t address() const { return memory ? memory->address + offset : 0; }
};

struct ImageView {
    Image* image = nullptr;
    int32_t format = 0;
    uint32_t base_level = 0, level_count = 1;
};

struct Sampler {
    cpv_sampler_desc desc{};
};

struct ShaderModule {
    std::vector<uint32_t> words;
};

struct DescriptorSetLayout {
    std::map<uint32_t, int32_t> types;   // binding -> descriptor type
};

struct PipelineLayout {
    std::vector<DescriptorSetLayout*> sets;
};

struct Descriptor {
    int32_t type = -1;
    Buffer* buffer = nullptr;
    VkDeviceSize offset = 0, range = 0;
    ImageView* view = nullptr;
    Sampler* sampler = nullptr;
};

struct DescriptorSet {
    std::map<uint32_t, Descriptor> bindings;
};

struct DescriptorPool {
    std::vector<DescriptorSet*> sets;
};

struct RenderPass {
    std::vector<VkAttachmentDescription> attachments;
    int32_t color = -1, depth = -1;   // attachment indices used by the (single) subpass
};

struct Framebuffer {
    RenderPass* pass = nullptr;
    std::vector<ImageView*> views;
    uint32_t width = 0, height = 0;
};

struct Pipeline {
    cpv_pipeline_desc desc{};
    bool dynamic_viewport = false;
    VkViewport viewport{};
    // state taken from vkCmdSet* at draw time instead of from the pipeline
    bool dyn_blend_constants = false, dyn_depth_bounds = false;
    bool dyn_compare_mask = false, dyn_write_mask = false, dyn_reference = false;
};

struct CommandPool {};

enum class Op {
    BeginPass, EndPass, BindPipeline, BindVertex, BindIndex, BindSets, SetViewport, Draw, DrawIndexed, DrawIndirect,
    DrawIndexedIndirect, DrawIndirectCount, DrawIndexedIndirectCount, CopyBuffer, CopyImage, CopyBufferToImage, CopyImageToBuffer, FillBuffer, UpdateBuffer, ClearColorImage,
    ClearDepthStencilImage, ClearAttachments, BlitImage, SetBlendConstants, SetDepthBounds, SetStencilCompareMask,
    SetStencilWriteMask, SetStencilReference
};

struct Cmd {
    Op op;
    RenderPass* pass = nullptr;
    Framebuffer* fb = nullptr;
    std::vector<VkClearValue> clears;
    Pipeline* pipeline = nullptr;
    uint32_t first = 0;
    std::vector<Buffer*> buffers;
    std::vector<VkDeviceSize> offsets;
    std::vector<DescriptorSet*> sets;
    int32_t index_type = 0;
    VkViewport viewport{};
    uint32_t count = 0, instances = 0, first_vertex = 0, first_instance = 0;
    int32_t vertex_offset = 0;
    // transfers, clears and indirect draws
    Buffer* src_buffer = nullptr;
    Buffer* count_buffer = nullptr;     // DrawIndirectCount: the draw count's buffer, at count_offset
    VkDeviceSize count_offset = 0;
    Buffer* dst_buffer = nullptr;
    Image* src_image = nullptr;
    Image* dst_image = nullptr;
    std::vector<VkBufferCopy> buffer_copies;
    std::vector<VkImageCopy> image_copies;
    std::vector<VkBufferImageCopy> buffer_image_copies;
    std::vector<VkImageBlit> blits;
    std::vector<VkClearAttachment> attachments;
    std::vector<VkClearRect> rects;
    std::vector<VkImageSubresourceRange> ranges;
    VkClearValue value{};
    VkDeviceSize offset = 0, size = 0;
    uint32_t data = 0, stride = 0;
    int32_t filter = 0;
    std::vector<uint8_t> bytes;
    float values[4] = {};   // dynamic state: blend constants, depth bounds (min, max)
};

struct CommandBuffer : Dispatchable {
    Device* device = nullptr;
    std::vector<Cmd> cmds;
    bool one_time = false;          // recorded with VK_COMMAND_BUFFER_USAGE_ONE_TIME_SUBMIT_BIT
    cpv_batch* batch = nullptr;     // its render passes recorded once, replayed by every submit
    bool batch_tried = false;       // a batch was attempted (and failed or was not possible)
    void forget()
    {
        if (batch) cpv_batch_destroy(batch);
        batch = nullptr;
        batch_tried = false;
        cmds.clear();
    }
    ~CommandBuffer() { forget(); }
};

// A VkFence: a libcpv fence (a HIP event recorded after the submit's work).
struct Fence {
    cpv_fence* f = nullptr;
};

template <class H, class T> H handle(T* p) { return reinterpret_cast<H>(p); }

VkResult device_error(int rc)
{
    if (rc == CPV_OK) return VK_SUCCESS;
    if (rc == CPV_ERROR_OUT_OF_MEMORY) return VK_ERROR_OUT_OF_DEVICE_MEMORY;
    if (rc == CPV_ERROR_UNSUPPORTED) return VK_ERROR_FEATURE_NOT_PRESENT;
    if (rc == CPV_ERROR_COMPILE) return VK_ERROR_INITIALIZATION_FAILED;
    std::fprintf(stderr, "cpvulkan_amd: %s\n", cpv_last_error());
    return VK_ERROR_DEVICE_LOST;
}

// -------------------------------------------------------- instance / device

VkResult CreateInstance(const VkInstanceCreateInfo* info, const VkAllocationCallbacks*, VkInstance* out)
{
    if (info && info->enabledExtensionCount) return VK_ERROR_EXTENSION_NOT_PRESENT;
    if (info && info->enabledLayerCount) return VK_ERROR_LAYER_NOT_PRESENT;
    Instance* i = new (std::nothrow) Instance();
    if (!i) return VK_ERROR_OUT_OF_HOST_MEMORY;
    i->physical.instance = i;
    *out = reinterpret_cast<VkInstance>(i);
    return VK_SUCCESS;
}

void DestroyInstance(VkInstance inst, const VkAllocationCallbacks*) { delete reinterpret_cast<Instance*>(inst); }

VkResult EnumeratePhysicalDevices(VkInstance inst, uint32_t* count, VkPhysicalDevice* out)
{
    if (!out) { *count = 1; return VK_SUCCESS; }
    if (*count < 1) return VK_INCOMPLETE;
    *count = 1;
    out[0] = reinterpret_cast<VkPhysicalDevice>(&reinterpret_cast<Instance*>(inst)->physical);
    return VK_SUCCESS;
}

void GetPhysicalDeviceProperties(VkPhysicalDevice, VkPhysicalDeviceProperties* p)
{
    std::memset(p, 0, sizeof(*p));
    p->apiVersion = VK_API_VERSION_1_1;
    p->driverVersion = VK_MAKE_VERSION(0, 1, 0);
    p->vendorID = 0x1002;
    p->deviceType = VK_PHYSICAL_DEVICE_TYPE_DISCRETE_GPU;
    std::snprintf(p->deviceName, sizeof(p->deviceName), "cpvulkan_amd draw path on MI355X (gfx950)");
    VkPhysicalDeviceLimits& l = p->limits;
    l.maxImageDimension1D = l.maxImageDimension2D = 16384;
    l.maxImageDimension3D = l.maxImageDimensionCube = 1;
    l.maxImageArrayLayers = 1;
    l.maxUniformBufferRange = 65536;
    l.maxMemoryAllocationCount = 1u << 20;
    l.maxSamplerAllocationCount = 4000;
    l.bufferImageGranularity = 1;
    l.maxBoundDescriptorSets = 1;
    l.maxPerStageDescriptorSamplers = l.maxPerStageDescriptorUniformBuffers = l.maxPerStageDescriptorSampledImages = 4;
    l.maxPerStageResources = 8;
    l.maxDescriptorSetSamplers = l.maxDescriptorSetUniformBuffers = l.maxDescriptorSetSampledImages = 4;
    l.maxVertexInputAttributes = l.maxVertexInputBindings = 16;
    l.maxVertexInputAttributeOffset = 2047;
    l.maxVertexInputBindingStride = 2048;
    l.maxVertexOutputComponents = l.maxFragmentInputComponents = 64;
    l.maxFragmentOutputAttachments = 1;
    l.maxFragmentCombinedOutputResources = 8;
    l.maxDrawIndexedIndexValue = 0xFFFFFFFFu;
    l.maxSamplerLodBias = 32.0f;
    l.maxSamplerAnisotropy = 1.0f;
    l.maxViewports = 1;
    l.maxViewportDimensions[0] = l.maxViewportDimensions[1] = 16384;
    l.viewportBoundsRange[0] = -32768.0f;
    l.viewportBoundsRange[1] = 32767.0f;
    l.minMemoryMapAlignment = 64;
    l.minTexelBufferOffsetAlignment = l.minUniformBufferOffsetAlignment = l.minStorageBufferOffsetAlignment = 16;
    l.maxFramebufferWidth = l.maxFramebufferHeight = 16384;
    l.maxFramebufferLayers = 1;
    l.framebufferColorSampleCounts = l.framebufferDepthSampleCounts = l.framebufferStencilSampleCounts = 1;
    l.framebufferNoAttachmentsSampleCounts = 1;
    l.maxColorAttachments = 1;
    l.sampledImageColorSampleCounts = l.sampledImageIntegerSampleCounts = l.sampledImageDepthSampleCounts = 1;
    l.sampledImageStencilSampleCounts = l.storageImageSampleCounts = 1;
    l.maxSampleMaskWords = 1;
    l.timestampPeriod = 1.0f;
    l.discreteQueuePriorities = 2;
    // CPVulkanBase/Config.h:192-195: wide lines and large points, granularity 1
    l.pointSizeRange[0] = 1.0f;
    l.pointSizeRange[1] = 64.0f - 1.0f / 32.0f;
    l.lineWidthRange[0] = 1.0f;
    l.lineWidthRange[1] = 8.0f - 1.0f / 32.0f;
    l.pointSizeGranularity = l.lineWidthGranularity = 1.0f;
    l.optimalBufferCopyOffsetAlignment = l.optimalBufferCopyRowPitchAlignment = l.nonCoherentAtomSize = 1;
}

// VkPhysicalDeviceFeatures in declaration order: depthBounds (14), wideLines
// (15) and largePoints (16) are on, as in PhysicalDevice.cpp:214-216.
void GetPhysicalDeviceFeatures(VkPhysicalDevice, VkPhysicalDeviceFeatures* f)
{
    std::memset(f, 0, sizeof(*f));
    f->feature[14] = f->feature[15] = f->feature[16] = 1;
}

void GetPhysicalDeviceMemoryProperties(VkPhysicalDevice, VkPhysicalDeviceMemoryProperties* m)
{
    std::memset(m, 0, sizeof(*m));
    m->memoryTypeCount = 2;
    m->memoryTypes[0] = VkMemoryType{VK_MEMORY_PROPERTY_DEVICE_LOCAL_BIT, 0};
    m->memoryTypes[1] = VkMemoryType{VK_MEMORY_PROPERTY_HOST_VISIBLE_BIT | VK_MEMORY_PROPERTY_HOST_COHERENT_BIT, 1};
    m->memoryHeapCount = 2;
    m->memoryHeaps[0] = VkMemoryHeap{288ull << 30, VK_MEMORY_HEAP_DEVICE_LOCAL_BIT};
    m->memoryHeaps[1] = VkMemoryHeap{64ull << 30, 0};
}

void GetPhysicalDeviceQueueFamilyProperties(VkPhysicalDevice, uint32_t* count, VkQueueFamilyProperties* out)
{
    if (!out) { *count = 1; return; }
    if (*count < 1) return;
    *count = 1;
    out[0] = VkQueueFamilyProperties{VK_QUEUE_GRAPHICS_BIT | VK_QUEUE_TRANSFER_BIT, 1, 64, {1, 1, 1}};
}

void GetPhysicalDeviceFormatProperties(VkPhysicalDevice, int32_t format, VkFormatProperties* p)
{
    std::memset(p, 0, sizeof(*p));
    if (format_bytes(format)) {
        // sampled image | colour attachment | depth/stencil attachment | vertex buffer
        p->linearTilingFeatures = p->optimalTilingFeatures = is_depth_stencil(format) ? 0x200u : (0x1u | 0x80u);
        p->bufferFeatures = is_depth_stencil(format) ? 0u : 0x40u;
    }
}

VkResult EnumerateExtensionProperties(const char*, uint32_t* count, VkExtensionProperties*)
{
    *count = 0;
    return VK_SUCCESS;
}

// Device extensions: the indirect-count draws (CPVulkan/Extensions.cpp:765-775, 1030-1040).
const VkExtensionProperties kDeviceExtensions[] = {{"VK_KHR_draw_indirect_count", 1}, {"VK_AMD_draw_indirect_count", 2}};

VkResult EnumerateDeviceExtensionProperties(VkPhysicalDevice, const char* layer, uint32_t* count, VkExtensionProperties* out)
{
    const uint32_t n = layer ? 0u : (uint32_t)(sizeof(kDeviceExtensions) / sizeof(kDeviceExtensions[0]));
    if (!out) {
        *count = n;
        return VK_SUCCESS;
    }
    const uint32_t k = std::min(*count, n);
    for (uint32_t i = 0; i < k; i++) out[i] = kDeviceExtensions[i];
    *count = k;
    return k < n ? VK_INCOMPLETE : VK_SUCCESS;
}

VkResult EnumerateLayerProperties(uint32_t* count, VkLayerProperties*)
{
    *count = 0;
    return VK_SUCCESS;
}

VkResult CreateDevice(VkPhysicalDevice, const VkDeviceCreateInfo* info, const VkAllocationCallbacks*, VkDevice* out)
{
    for (uint32_t i = 0; info && i < info->enabledExtensionCount; i++) {
        bool known = false;
        for (const VkExtensionProperties& e : kDeviceExtensions) known = known || std::strcmp(e.extensionName, info->ppEnabledExtensionNames[i]) == 0;
        if (!known) return VK_ERROR_EXTENSION_NOT_PRESENT;
    }
    Device* d = new (std::nothrow) Device();
    if (!d) return VK_ERROR_OUT_OF_HOST_MEMORY;
    const int rc = cpv_context_create(0, &d->ctx);
    if (rc) {
        delete d;
        return VK_ERROR_INITIALIZATION_FAILED;
    }
    d->queue.device = d;
    *out = reinterpret_cast<VkDevice>(d);
    return VK_SUCCESS;
}

void DestroyDevice(VkDevice dev, const VkAllocationCallbacks*)
{
    Device* d = reinterpret_cast<Device*>(dev);
    if (!d) return;
    cpv_context_destroy(d->ctx);
    delete d;
}

void GetDeviceQueue(VkDevice dev, uint32_t, uint32_t, VkQueue* out)
{
    *out = reinterpret_cast<VkQueue>(&reinterpret_cast<Device*>(dev)->queue);
}

VkResult DeviceWaitIdle(VkDevice dev) { return device_error(cpv_synchronize(reinterpret_cast<Device*>(dev)->ctx)); }

// ---------------------------------------------------------------- memory

VkResult AllocateMemory(VkDevice dev, const VkMemoryAllocateInfo* info, const VkAllocationCallbacks*, VkDeviceMemory* out)
{
    Device* d = reinterpret_cast<Device*>(dev);
    if (info->memoryTypeIndex > 1) return VK_ERROR_OUT_OF_DEVICE_MEMORY;
    Memory* m = new Memory();
    m->size = info->allocationSize;
    m->host = info->memoryTypeIndex == 1;
    const bool host = m->host;
    const int rc = host ? cpv_alloc_host(d->ctx, m->size, &m->address) : cpv_alloc(d->ctx, m->size, &m->address);
    if (rc) {
        delete m;
        return host ? VK_ERROR_OUT_OF_HOST_MEMORY : VK_ERROR_OUT_OF_DEVICE_MEMORY;
    }
    *out = handle<VkDeviceMemory>(m);
    return VK_SUCCESS;
}

void FreeMemory(VkDevice dev, VkDeviceMemory mem, const VkAllocationCallbacks*)
{
    Memory* m = reinterpret_cast<Memory*>(mem);
    if (!m) return;
    Device* d = reinterpret_cast<Device*>(dev);
    if (m->host) cpv_free_host(d->ctx, m->address);
    else cpv_free(d->ctx, m->address);
    delete m;
}

VkResult MapMemory(VkDevice, VkDeviceMemory mem, VkDeviceSize offset, VkDeviceSize, VkFlags, void** data)
{
    Memory* m = reinterpret_cast<Memory*>(mem);
    if (!m->host) return VK_ERROR_MEMORY_MAP_FAILED;
    *data = reinterpret_cast<void*>(m->address + offset);
    return VK_SUCCESS;
}

void UnmapMemory(VkDevice, VkDeviceMemory) {}
VkResult FlushMappedMemoryRanges(VkDevice, uint32_t, const VkMappedMemoryRange*) { return VK_SUCCESS; }
VkResult InvalidateMappedMemoryRanges(VkDevice, uint32_t, const VkMappedMemoryRange*) { return VK_SUCCESS; }

VkResult CreateBuffer(VkDevice, const VkBufferCreateInfo* info, const VkAllocationCallbacks*, VkBuffer* out)
{
    Buffer* b = new Buffer();
    b->size = info->size;
    *out = handle<VkBuffer>(b);
    return VK_SUCCESS;
}

void DestroyBuffer(VkDevice, VkBuffer b, const VkAllocationCallbacks*) { delete reinterpret_cast<Buffer*>(b); }

void GetBufferMemoryRequirements(VkDevice, VkBuffer b, VkMemoryRequirements* r)
{
    r->size = (reinterpret_cast<Buffer*>(b)->size + 255) & ~255ull;
    r->alignment = 256;
    r->memoryTypeBits = 3;
}

VkResult BindBufferMemory(VkDevice, VkBuffer b, VkDeviceMemory mem, VkDeviceSize offset)
{
    Buffer* buf = reinterpret_cast<Buffer*>(b);
    buf->memory = reinterpret_cast<Memory*>(mem);
    buf->offset = offset;
    return VK_SUCCESS;
}

VkResult CreateImage(VkDevice, const VkImageCreateInfo* info, const VkAllocationCallbacks*, VkImage* out)
{
    const uint32_t px = format_bytes(info->format);
    if (!px || info->imageType != VK_IMAGE_TYPE_2D || info->extent.depth != 1 || info->arrayLayers != 1 || info->samples != 1)
        return VK_ERROR_FORMAT_NOT_SUPPORTED;
    Image* im = new Image();
    im->format = info->format;
    im->width = info->extent.width;
    im->height = info->extent.height;
    im->levels = info->mipLevels ? info->mipLevels : 1;
    // GetNormalImageSize (CPVulkanBase/Formats.cpp:455-482): levels packed one after another
    uint32_t w = im->width, h = im->height;
    for (uint32_t l = 0; l < im->levels; l++) {
        im->level_offset.push_back(im->size);
        im->size += (VkDeviceSize)w * h * px;
        w = w > 1 ? w / 2 : 1;
        h = h > 1 ? h / 2 : 1;
    }
    *out = handle<VkImage>(im);
    return VK_SUCCESS;
}

void DestroyImage(VkDevice, VkImage im, const VkAllocationCallbacks*) { delete reinterpret_cast<Image*>(im); }

void GetImageMemoryRequirements(VkDevice, VkImage im, VkMemoryRequirements* r)
{
    r->size = (reinterpret_cast<Image*>(im)->size + 255) & ~255ull;
    r->alignment = 256;
    r->memoryTypeBits = 3;
}

VkResult BindImageMemory(VkDevice, VkImage im, VkDeviceMemory mem, VkDeviceSize offset)
{
    Image* i = reinterpret_cast<Image*>(im);
    i->memory = reinterpret_cast<Memory*>(mem);
    i->offset = offset;
    return VK_SUCCESS;
}

void GetImageSubresourceLayout(VkDevice, VkImage im, const VkImageSubresource* sub, VkSubresourceLayout* out)
{
    const Image* i = reinterpret_cast<Image*>(im);
    const uint32_t px = format_bytes(i->format);
    const uint32_t l = sub->mipLevel < i->levels ? sub->mipLevel : 0;
    const uint32_t w = std::max(i->width >> l, 1u), h = std::max(i->height >> l, 1u);
    out->offset = i->level_offset[l];
    out->rowPitch = (VkDeviceSize)w * px;
    out->size = out->rowPitch * h;
    out->arrayPitch = i->size;
    out->depthPitch = out->size;
}

VkResult CreateImageView(VkDevice, const VkImageViewCreateInfo* info, const VkAllocationCallbacks*, VkImageView* out)
{
    ImageView* v = new ImageView();
    v->image = reinterpret_cast<Image*>(info->image);
    v->format = info->format;
    v->base_level = info->subresourceRange.baseMipLevel;
    v->level_count = info->subresourceRange.levelCount == ~0u ? v->image->levels - v->base_level : info->subresourceRange.levelCount;
    *out = handle<VkImageView>(v);
    return VK_SUCCESS;
}

void DestroyImageView(VkDevice, VkImageView v, const VkAllocationCallbacks*) { delete reinterpret_cast<ImageView*>(v); }

VkResult CreateSampler(VkDevice, const VkSamplerCreateInfo* info, const VkAllocationCallbacks*, VkSampler* out)
{
    if (info->anisotropyEnable || info->compareEnable) return VK_ERROR_FEATURE_NOT_PRESENT;
    Sampler* s = new Sampler();
    cpv_sampler_desc& d = s->desc;
    d.mag_filter = (uint32_t)info->magFilter;
    d.min_filter = (uint32_t)info->minFilter;
    d.mipmap_mode = (uint32_t)info->mipmapMode;
    d.address_u = (uint32_t)info->addressModeU;
    d.address_v = (uint32_t)info->addressModeV;
    d.address_w = (uint32_t)info->addressModeW;
    d.mip_lod_bias = info->mipLodBias;
    d.min_lod = info->minLod;
    d.max_lod = info->maxLod;
    d.border_color = (uint32_t)info->borderColor;
    d.unnormalized = info->unnormalizedCoordinates;
    *out = handle<VkSampler>(s);
    return VK_SUCCESS;
}

void DestroySampler(VkDevice, VkSampler s, const VkAllocationCallbacks*) { delete reinterpret_cast<Sampler*>(s); }

VkResult CreateShaderModule(VkDevice, const VkShaderModuleCreateInfo* info, const VkAllocationCallbacks*, VkShaderModule* out)
{
    ShaderModule* m = new ShaderModule();
    m->words.assign(info->pCode, info->pCode + info->codeSize / 4);
    *out = handle<VkShaderModule>(m);
    return VK_SUCCESS;
}

void DestroyShaderModule(VkDevice, VkShaderModule m, const VkAllocationCallbacks*) { delete reinterpret_cast<ShaderModule*>(m); }

// --------------------------------------------------------------- descriptors

VkResult CreateDescriptorSetLayout(VkDevice, const VkDescriptorSetLayoutCreateInfo* info, const VkAllocationCallbacks*,
                                   VkDescriptorSetLayout* out)
{
    DescriptorSetLayout* l = new DescriptorSetLayout();
    for (uint32_t i = 0; i < info->bindingCount; i++) {
        const VkDescriptorSetLayoutBinding& b = info->pBindings[i];
        if (b.descriptorType != VK_DESCRIPTOR_TYPE_UNIFORM_BUFFER && b.descriptorType != VK_DESCRIPTOR_TYPE_COMBINED_IMAGE_SAMPLER) {
            delete l;
            return VK_ERROR_FEATURE_NOT_PRESENT;
        }
        l->types[b.binding] = b.descriptorType;
    }
    *out = handle<VkDescriptorSetLayout>(l);
    return VK_SUCCESS;
}

void DestroyDescriptorSetLayout(VkDevice, VkDescriptorSetLayout l, const VkAllocationCallbacks*)
{
    delete reinterpret_cast<DescriptorSetLayout*>(l);
}

VkResult CreatePipelineLayout(VkDevice, const VkPipelineLayoutCreateInfo* info, const VkAllocationCallbacks*, VkPipelineLayout* out)
{
    PipelineLayout* l = new PipelineLayout();
    for (uint32_t i = 0; i < info->setLayoutCount; i++) l->sets.push_back(reinterpret_cast<DescriptorSetLayout*>(info->pSetLayouts[i]));
    *out = handle<VkPipelineLayout>(l);
    return VK_SUCCESS;
}

void DestroyPipelineLayout(VkDevice, VkPipelineLayout l, const VkAllocationCallbacks*) { delete reinterpret_cast<PipelineLayout*>(l); }

VkResult CreateDescriptorPool(VkDevice, const VkDescriptorPoolCreateInfo*, const VkAllocationCallbacks*, VkDescriptorPool* out)
{
    *out = handle<VkDescriptorPool>(new DescriptorPool());
    return VK_SUCCESS;
}

void DestroyDescriptorPool(VkDevice, VkDescriptorPool p, const VkAllocationCallbacks*)
{
    DescriptorPool* pool = reinterpret_cast<DescriptorPool*>(p);
    if (!pool) return;
    for (DescriptorSet* s : pool->sets) delete s;
    delete pool;
}

VkResult AllocateDescriptorSets(VkDevice, const VkDescriptorSetAllocateInfo* info, VkDescriptorSet* out)
{
    DescriptorPool* pool = reinterpret_cast<DescriptorPool*>(info->descriptorPool);
    for (uint32_t i = 0; i < info->descriptorSetCount; i++) {
        DescriptorSet* s = new DescriptorSet();
        pool->sets.push_back(s);
        out[i] = handle<VkDescriptorSet>(s);
    }
    return VK_SUCCESS;
}

VkResult FreeDescriptorSets(VkDevice, VkDescriptorPool p, uint32_t n, const VkDescriptorSet* sets)
{
    DescriptorPool* pool = reinterpret_cast<DescriptorPool*>(p);
    for (uint32_t i = 0; i < n; i++) {
        DescriptorSet* s = reinterpret_cast<DescriptorSet*>(sets[i]);
        for (size_t k = 0; k < pool->sets.size(); k++)
            if (pool->sets[k] == s) { pool->sets.erase(pool->sets.begin() + (long)k); break; }
        delete s;
    }
    return VK_SUCCESS;
}

void UpdateDescriptorSets(VkDevice, uint32_t n, const VkWriteDescriptorSet* writes, uint32_t copies, const VkCopyDescriptorSet*)
{
    (void)copies;
    for (uint32_t i = 0; i < n; i++) {
        const VkWriteDescriptorSet& w = writes[i];
        DescriptorSet* s = reinterpret_cast<DescriptorSet*>(w.dstSet);
        for (uint32_t k = 0; k < w.descriptorCount; k++) {
            Descriptor& d = s->bindings[w.dstBinding + w.dstArrayElement + k];
            d.type = w.descriptorType;
            if (w.descriptorType == VK_DESCRIPTOR_TYPE_UNIFORM_BUFFER && w.pBufferInfo) {
                d.buffer = reinterpret_cast<Buffer*>(w.pBufferInfo[k].buffer);
                d.offset = w.pBufferInfo[k].offset;
                d.range = w.pBufferInfo[k].range;
            } else if (w.descriptorType == VK_DESCRIPTOR_TYPE_COMBINED_IMAGE_SAMPLER && w.pImageInfo) {
                d.view = reinterpret_cast<ImageView*>(w.pImageInfo[k].imageView);
                d.sampler = reinterpret_cast<Sampler*>(w.pImageInfo[k].sampler);
            }
        }
    }
}

// ------------------------------------------------------ passes and pipelines

VkResult CreateRenderPass(VkDevice, const VkRenderPassCreateInfo* info, const VkAllocationCallbacks*, VkRenderPass* out)
{
    if (info->subpassCount != 1) return VK_ERROR_FEATURE_NOT_PRESENT;
    const VkSubpassDescription& sp = info->pSubpasses[0];
    if (sp.colorAttachmentCount > 1 || sp.inputAttachmentCount || sp.pResolveAttachments) return VK_ERROR_FEATURE_NOT_PRESENT;
    RenderPass* rp = new RenderPass();
    rp->attachments.assign(info->pAttachments, info->pAttachments + info->attachmentCount);
    if (sp.colorAttachmentCount && sp.pColorAttachments[0].attachment != VK_ATTACHMENT_UNUSED)
        rp->color = (int32_t)sp.pColorAttachments[0].attachment;
    if (sp.pDepthStencilAttachment && sp.pDepthStencilAttachment->attachment != VK_ATTACHMENT_UNUSED)
        rp->depth = (int32_t)sp.pDepthStencilAttachment->attachment;
    *out = handle<VkRenderPass>(rp);
    return VK_SUCCESS;
}

void DestroyRenderPass(VkDevice, VkRenderPass rp, const VkAllocationCallbacks*) { delete reinterpret_cast<RenderPass*>(rp); }

VkResult CreateFramebuffer(VkDevice, const VkFramebufferCreateInfo* info, const VkAllocationCallbacks*, VkFramebuffer* out)
{
    Framebuffer* fb = new Framebuffer();
    fb->pass = reinterpret_cast<RenderPass*>(info->renderPass);
    for (uint32_t i = 0; i < info->attachmentCount; i++) fb->views.push_back(reinterpret_cast<ImageView*>(info->pAttachments[i]));
    fb->width = info->width;
    fb->height = info->height;
    *out = handle<VkFramebuffer>(fb);
    return VK_SUCCESS;
}

void DestroyFramebuffer(VkDevice, VkFramebuffer fb, const VkAllocationCallbacks*) { delete reinterpret_cast<Framebuffer*>(fb); }

VkResult build_pipeline(Device* d, const VkGraphicsPipelineCreateInfo& ci, Pipeline* p)
{
    cpv_pipeline_desc& o = p->desc;
    std::memset(&o, 0, sizeof(o));
    const ShaderModule* vs = nullptr;
    const ShaderModule* fs = nullptr;
    const char* vs_entry = "main";
    const char* fs_entry = "main";
    for (uint32_t i = 0; i < ci.stageCount; i++) {
        const VkPipelineShaderStageCreateInfo& s = ci.pStages[i];
        if (s.stage == VK_SHADER_STAGE_VERTEX_BIT) { vs = reinterpret_cast<ShaderModule*>(s.module); vs_entry = s.pName; }
        else if (s.stage == VK_SHADER_STAGE_FRAGMENT_BIT) { fs = reinterpret_cast<ShaderModule*>(s.module); fs_entry = s.pName; }
        else return VK_ERROR_FEATURE_NOT_PRESENT;
    }
    if (!vs || !fs) return VK_ERROR_FEATURE_NOT_PRESENT;
    // the JIT step of GraphicsPipeline::Create (CPVulkan/Pipeline.cpp:599-714): SPIR-V -> HIP -> hiprtc
    const int rc = cpv_program_create_spirv(d->ctx, vs->words.data(), vs->words.size(), vs_entry, fs->words.data(), fs->words.size(),
                                            fs_entry, &o.program);
    if (rc) {
        std::fprintf(stderr, "cpvulkan_amd: pipeline: %s\n", cpv_last_error());
        return rc == CPV_ERROR_UNSUPPORTED ? VK_ERROR_FEATURE_NOT_PRESENT : VK_ERROR_INITIALIZATION_FAILED;
    }
    if (const VkPipelineVertexInputStateCreateInfo* vi = ci.pVertexInputState) {
        if (vi->vertexBindingDescriptionCount > CPV_MAX_VERTEX_BINDINGS || vi->vertexAttributeDescriptionCount > CPV_MAX_VERTEX_ATTRIBUTES)
            return VK_ERROR_FEATURE_NOT_PRESENT;
        o.binding_count = vi->vertexBindingDescriptionCount;
        for (uint32_t i = 0; i < o.binding_count; i++) {
            const VkVertexInputBindingDescription& b = vi->pVertexBindingDescriptions[i];
            o.bindings[i] = cpv_vertex_binding_desc{b.binding, b.stride, (uint32_t)b.inputRate, 0};
        }
        o.attribute_count = vi->vertexAttributeDescriptionCount;
        for (uint32_t i = 0; i < o.attribute_count; i++) {
            const VkVertexInputAttributeDescription& a = vi->pVertexAttributeDescriptions[i];
            o.attributes[i] = cpv_vertex_attribute_desc{a.location, a.binding, (uint32_t)a.format, a.offset};
        }
    }
    if (ci.pInputAssemblyState) {
        o.topology = (uint32_t)ci.pInputAssemblyState->topology;
        o.primitive_restart = ci.pInputAssemblyState->primitiveRestartEnable;
    }
    o.line_width = 1.0f;
    if (const VkPipelineRasterizationStateCreateInfo* r = ci.pRasterizationState) {
        o.line_width = r->lineWidth;   // RasterizationState.LineWidth (Pipeline.cpp:224); ProcessLines reads no dynamic width
        o.cull_mode = r->cullMode;
        o.front_face = (uint32_t)r->frontFace;
        o.polygon_mode = (uint32_t)r->polygonMode;
        o.rasterizer_discard = r->rasterizerDiscardEnable;
        o.depth_clamp = r->depthClampEnable;
    }
    if (const VkPipelineDepthStencilStateCreateInfo* ds = ci.pDepthStencilState) {
        o.depth_test = ds->depthTestEnable;
        o.depth_write = ds->depthWriteEnable;
        o.depth_compare = (uint32_t)ds->depthCompareOp;
        o.depth_bounds_test = ds->depthBoundsTestEnable;
        o.stencil_test = ds->stencilTestEnable;
        o.min_depth_bounds = ds->minDepthBounds;
        o.max_depth_bounds = ds->maxDepthBounds;
        const VkStencilOpState* st[2] = {&ds->front, &ds->back};
        cpv_stencil_state* dst[2] = {&o.front, &o.back};
        for (int k = 0; k < 2; k++)
            *dst[k] = cpv_stencil_state{(uint32_t)st[k]->failOp, (uint32_t)st[k]->passOp, (uint32_t)st[k]->depthFailOp,
                                        (uint32_t)st[k]->compareOp, st[k]->compareMask, st[k]->writeMask, st[k]->reference, 0};
    }
    RenderPass* rp = reinterpret_cast<RenderPass*>(ci.renderPass);
    if (rp && rp->color >= 0) {
        o.color_count = 1;
        o.color_formats[0] = (uint32_t)rp->attachments[rp->color].format;
    }
    if (rp && rp->depth >= 0) o.depth_stencil_format = (uint32_t)rp->attachments[rp->depth].format;
    o.blend[0].color_write_mask = CPV_COLOR_RGBA;
    if (const VkPipelineColorBlendStateCreateInfo* cb = ci.pColorBlendState) {
        o.logic_op_enable = cb->logicOpEnable;
        for (int k = 0; k < 4; k++) o.blend_constants[k] = cb->blendConstants[k];
        if (cb->attachmentCount) {
            const VkPipelineColorBlendAttachmentState& a = cb->pAttachments[0];
            o.blend[0] = cpv_blend_attachment{a.blendEnable, (uint32_t)a.srcColorBlendFactor, (uint32_t)a.dstColorBlendFactor,
                                              (uint32_t)a.colorBlendOp, (uint32_t)a.srcAlphaBlendFactor, (uint32_t)a.dstAlphaBlendFactor,
                                              (uint32_t)a.alphaBlendOp, a.colorWriteMask};
        }
    }
    o.origin_upper_left = 1;   // Vulkan fragment shaders declare OriginUpperLeft (the translator requires it)
    if (ci.pDynamicState)
        for (uint32_t i = 0; i < ci.pDynamicState->dynamicStateCount; i++) {
            const int32_t ds = ci.pDynamicState->pDynamicStates[i];
            // The reference compiles these from the pipeline and stops
            // (TODO_ERROR) when they are dynamic and used
            // (PipelineCompiler.cpp:1141-1148, 1201-1214, 1305-1311); they take
            // the vkCmdSet* values current at each draw here.
            if (ds == VK_DYNAMIC_STATE_VIEWPORT) p->dynamic_viewport = true;
            if (ds == VK_DYNAMIC_STATE_BLEND_CONSTANTS) p->dyn_blend_constants = true;
            if (ds == VK_DYNAMIC_STATE_DEPTH_BOUNDS) p->dyn_depth_bounds = true;
            if (ds == VK_DYNAMIC_STATE_STENCIL_COMPARE_MASK) p->dyn_compare_mask = true;
            if (ds == VK_DYNAMIC_STATE_STENCIL_WRITE_MASK) p->dyn_write_mask = true;
            if (ds == VK_DYNAMIC_STATE_STENCIL_REFERENCE) p->dyn_reference = true;
        }
    if (!p->dynamic_viewport && ci.pViewportState && ci.pViewportState->viewportCount && ci.pViewportState->pViewports)
        p->viewport = ci.pViewportState->pViewports[0];
    return VK_SUCCESS;
}

VkResult CreateGraphicsPipelines(VkDevice dev, VkPipelineCache, uint32_t n, const VkGraphicsPipelineCreateInfo* infos,
                                 const VkAllocationCallbacks*, VkPipeline* out)
{
    Device* d = reinterpret_cast<Device*>(dev);
    VkResult result = VK_SUCCESS;
    for (uint32_t i = 0; i < n; i++) {
        Pipeline* p = new Pipeline();
        const VkResult r = build_pipeline(d, infos[i], p);
        if (r != VK_SUCCESS) {
            if (p->desc.program) cpv_program_destroy(d->ctx, p->desc.program);
            delete p;
            out[i] = nullptr;
            result = r;
            continue;
        }
        out[i] = handle<VkPipeline>(p);
    }
    return result;
}

void DestroyPipeline(VkDevice dev, VkPipeline pl, const VkAllocationCallbacks*)
{
    Pipeline* p = reinterpret_cast<Pipeline*>(pl);
    if (!p) return;
    Device* d = reinterpret_cast<Device*>(dev);
    cpv_synchronize(d->ctx);
    if (p->desc.program) cpv_program_destroy(d->ctx, p->desc.program);
    delete p;
}

// ------------------------------------------------------- command buffers

VkResult CreateCommandPool(VkDevice, const VkCommandPoolCreateInfo*, const VkAllocationCallbacks*, VkCommandPool* out)
{
    *out = handle<VkCommandPool>(new CommandPool());
    return VK_SUCCESS;
}

void DestroyCommandPool(VkDevice, VkCommandPool p, const VkAllocationCallbacks*) { delete reinterpret_cast<CommandPool*>(p); }

VkResult AllocateCommandBuffers(VkDevice dev, const VkCommandBufferAllocateInfo* info, VkCommandBuffer* out)
{
    for (uint32_t i = 0; i < info->commandBufferCount; i++) {
        CommandBuffer* cb = new CommandBuffer();
        cb->device = reinterpret_cast<Device*>(dev);
        out[i] = reinterpret_cast<VkCommandBuffer>(cb);
    }
    return VK_SUCCESS;
}

void FreeCommandBuffers(VkDevice, VkCommandPool, uint32_t n, const VkCommandBuffer* cbs)
{
    for (uint32_t i = 0; i < n; i++) delete reinterpret_cast<CommandBuffer*>(cbs[i]);
}

// Beginning a command buffer resets it (CommandBuffer.cpp:964-968): a batch
// recorded from its previous contents is dropped.
VkResult BeginCommandBuffer(VkCommandBuffer cb, const VkCommandBufferBeginInfo* info)
{
    CommandBuffer* c = reinterpret_cast<CommandBuffer*>(cb);
    c->forget();
    c->one_time = info && (info->flags & VK_COMMAND_BUFFER_USAGE_ONE_TIME_SUBMIT_BIT);
    return VK_SUCCESS;
}

VkResult EndCommandBuffer(VkCommandBuffer) { return VK_SUCCESS; }

VkResult ResetCommandBuffer(VkCommandBuffer cb, VkFlags)
{
    reinterpret_cast<CommandBuffer*>(cb)->forget();
    return VK_SUCCESS;
}

CommandBuffer* CB(VkCommandBuffer cb) { return reinterpret_cast<CommandBuffer*>(cb); }

void CmdBeginRenderPass(VkCommandBuffer cb, const VkRenderPassBeginInfo* info, int32_t)
{
    Cmd c{Op::BeginPass};
    c.pass = reinterpret_cast<RenderPass*>(info->renderPass);
    c.fb = reinterpret_cast<Framebuffer*>(info->framebuffer);
    c.clears.assign(info->pClearValues, info->pClearValues + info->clearValueCount);
    CB(cb)->cmds.push_back(c);
}

void CmdEndRenderPass(VkCommandBuffer cb) { CB(cb)->cmds.push_back(Cmd{Op::EndPass}); }

void CmdBindPipeline(VkCommandBuffer cb, int32_t, VkPipeline p)
{
    Cmd c{Op::BindPipeline};
    c.pipeline = reinterpret_cast<Pipeline*>(p);
    CB(cb)->cmds.push_back(c);
}

void CmdBindVertexBuffers(VkCommandBuffer cb, uint32_t first, uint32_t n, const VkBuffer* bufs, const VkDeviceSize* offs)
{
    Cmd c{Op::BindVertex};
    c.first = first;
    for (uint32_t i = 0; i < n; i++) {
        c.buffers.push_back(reinterpret_cast<Buffer*>(bufs[i]));
        c.offsets.push_back(offs[i]);
    }
    CB(cb)->cmds.push_back(c);
}

void CmdBindIndexBuffer(VkCommandBuffer cb, VkBuffer b, VkDeviceSize off, int32_t type)
{
    Cmd c{Op::BindIndex};
    c.buffers.push_back(reinterpret_cast<Buffer*>(b));
    c.offsets.push_back(off);
    c.index_type = type;
    CB(cb)->cmds.push_back(c);
}

void CmdBindDescriptorSets(VkCommandBuffer cb, int32_t, VkPipelineLayout, uint32_t first, uint32_t n, const VkDescriptorSet* sets,
                           uint32_t, const uint32_t*)
{
    Cmd c{Op::BindSets};
    c.first = first;
    for (uint32_t i = 0; i < n; i++) c.sets.push_back(reinterpret_cast<DescriptorSet*>(sets[i]));
    CB(cb)->cmds.push_back(c);
}

void CmdSetViewport(VkCommandBuffer cb, uint32_t first, uint32_t n, const VkViewport* vps)
{
    if (first != 0 || n == 0) return;
    Cmd c{Op::SetViewport};
    c.viewport = vps[0];
    CB(cb)->cmds.push_back(c);
}

void CmdSetScissor(VkCommandBuffer, uint32_t, uint32_t, const VkRect2D*) {}   // ignored, as in the reference
// Dynamic state (CommandBuffer.Binding.cpp:234-242 records the depth
// bounds the same way); a draw reads the values current when it replays.
void CmdSetDepthBounds(VkCommandBuffer cb, float mn, float mx)
{
    Cmd c{Op::SetDepthBounds};
    c.values[0] = mn;
    c.values[1] = mx;
    CB(cb)->cmds.push_back(c);
}

void CmdSetBlendConstants(VkCommandBuffer cb, const float constants[4])
{
    Cmd c{Op::SetBlendConstants};
    for (int k = 0; k < 4; k++) c.values[k] = constants[k];
    CB(cb)->cmds.push_back(c);
}

void set_stencil(VkCommandBuffer cb, Op op, VkFlags faces, uint32_t v)
{
    Cmd c{op};
    c.first = faces;
    c.data = v;
    CB(cb)->cmds.push_back(c);
}
void CmdSetStencilCompareMask(VkCommandBuffer cb, VkFlags faces, uint32_t v) { set_stencil(cb, Op::SetStencilCompareMask, faces, v); }
void CmdSetStencilWriteMask(VkCommandBuffer cb, VkFlags faces, uint32_t v) { set_stencil(cb, Op::SetStencilWriteMask, faces, v); }
void CmdSetStencilReference(VkCommandBuffer cb, VkFlags faces, uint32_t v) { set_stencil(cb, Op::SetStencilReference, faces, v); }

// ProcessLines reads the pipeline's LineWidth only (CommandBuffer.Draw.cpp:1421),
// so a dynamic width changes nothing, as in the reference.
void CmdSetLineWidth(VkCommandBuffer, float) {}

void CmdDraw(VkCommandBuffer cb, uint32_t vertexCount, uint32_t instanceCount, uint32_t firstVertex, uint32_t firstInstance)
{
    Cmd c{Op::Draw};
    c.count = vertexCount;
    c.instances = instanceCount;
    c.first_vertex = firstVertex;
    c.first_instance = firstInstance;
    CB(cb)->cmds.push_back(c);
}

void CmdDrawIndexed(VkCommandBuffer cb, uint32_t indexCount, uint32_t instanceCount, uint32_t firstIndex, int32_t vertexOffset,
                    uint32_t firstInstance)
{
    Cmd c{Op::DrawIndexed};
    c.count = indexCount;
    c.instances = instanceCount;
    c.first_vertex = firstIndex;
    c.vertex_offset = vertexOffset;
    c.first_instance = firstInstance;
    CB(cb)->cmds.push_back(c);
}

void CmdDrawIndirect(VkCommandBuffer cb, VkBuffer buffer, VkDeviceSize offset, uint32_t drawCount, uint32_t stride)
{
    Cmd c{Op::DrawIndirect};
    c.src_buffer = reinterpret_cast<Buffer*>(buffer);
    c.offset = offset;
    c.count = drawCount;
    c.stride = stride;
    CB(cb)->cmds.push_back(c);
}

void CmdDrawIndexedIndirect(VkCommandBuffer cb, VkBuffer buffer, VkDeviceSize offset, uint32_t drawCount, uint32_t stride)
{
    Cmd c{Op::DrawIndexedIndirect};
    c.src_buffer = reinterpret_cast<Buffer*>(buffer);
    c.offset = offset;
    c.count = drawCount;
    c.stride = stride;
    CB(cb)->cmds.push_back(c);
}

// vkCmdDraw(Indexed)IndirectCount{KHR,AMD}: at most maxDrawCount records,
// as many as the count buffer holds when the command runs
// (DrawIndirectCountCommand / DrawIndexedIndirectCountCommand,
// CommandBuffer.Draw.cpp:2226-2348).
void draw_indirect_count(VkCommandBuffer cb, Op op, VkBuffer buffer, VkDeviceSize offset, VkBuffer countBuffer,
                         VkDeviceSize countBufferOffset, uint32_t maxDrawCount, uint32_t stride)
{
    Cmd c{op};
    c.src_buffer = reinterpret_cast<Buffer*>(buffer);
    c.offset = offset;
    c.count_buffer = reinterpret_cast<Buffer*>(countBuffer);
    c.count_offset = countBufferOffset;
    c.count = maxDrawCount;
    c.stride = stride;
    CB(cb)->cmds.push_back(c);
}

void CmdDrawIndirectCount(VkCommandBuffer cb, VkBuffer buffer, VkDeviceSize offset, VkBuffer countBuffer, VkDeviceSize countBufferOffset,
                          uint32_t maxDrawCount, uint32_t stride)
{
    draw_indirect_count(cb, Op::DrawIndirectCount, buffer, offset, countBuffer, countBufferOffset, maxDrawCount, stride);
}

void CmdDrawIndexedIndirectCount(VkCommandBuffer cb, VkBuffer buffer, VkDeviceSize offset, VkBuffer countBuffer,
                                 VkDeviceSize countBufferOffset, uint32_t maxDrawCount, uint32_t stride)
{
    draw_indirect_count(cb, Op::DrawIndexedIndirectCount, buffer, offset, countBuffer, countBufferOffset, maxDrawCount, stride);
}

void CmdCopyBuffer(VkCommandBuffer cb, VkBuffer src, VkBuffer dst, uint32_t n, const VkBufferCopy* regions)
{
    Cmd c{Op::CopyBuffer};
    c.src_buffer = reinterpret_cast<Buffer*>(src);
    c.dst_buffer = reinterpret_cast<Buffer*>(dst);
    c.buffer_copies.assign(regions, regions + n);
    CB(cb)->cmds.push_back(c);
}

void CmdCopyImage(VkCommandBuffer cb, VkImage src, int32_t, VkImage dst, int32_t, uint32_t n, const VkImageCopy* regions)
{
    Cmd c{Op::CopyImage};
    c.src_image = reinterpret_cast<Image*>(src);
    c.dst_image = reinterpret_cast<Image*>(dst);
    c.image_copies.assign(regions, regions + n);
    CB(cb)->cmds.push_back(c);
}

void CmdCopyBufferToImage(VkCommandBuffer cb, VkBuffer src, VkImage dst, int32_t, uint32_t n, const VkBufferImageCopy* regions)
{
    Cmd c{Op::CopyBufferToImage};
    c.src_buffer = reinterpret_cast<Buffer*>(src);
    c.dst_image = reinterpret_cast<Image*>(dst);
    c.buffer_image_copies.assign(regions, regions + n);
    CB(cb)->cmds.push_back(c);
}

void CmdCopyImageToBuffer(VkCommandBuffer cb, VkImage src, int32_t, VkBuffer dst, uint32_t n, const VkBufferImageCopy* regions)
{
    Cmd c{Op::CopyImageToBuffer};
    c.src_image = reinterpret_cast<Image*>(src);
    c.dst_buffer = reinterpret_cast<Buffer*>(dst);
    c.buffer_image_copies.assign(regions, regions + n);
    CB(cb)->cmds.push_back(c);
}

void CmdFillBuffer(VkCommandBuffer cb, VkBuffer dst, VkDeviceSize offset, VkDeviceSize size, uint32_t data)
{
    Cmd c{Op::FillBuffer};
    c.dst_buffer = reinterpret_cast<Buffer*>(dst);
    c.offset = offset;
    // VK_WHOLE_SIZE: to the end of the buffer (FillBufferCommand, CommandBuffer.cpp:280-289)
    c.size = size == VK_WHOLE_SIZE ? c.dst_buffer->size - offset : size;
    c.data = data;
    CB(cb)->cmds.push_back(c);
}

void CmdUpdateBuffer(VkCommandBuffer cb, VkBuffer dst, VkDeviceSize offset, VkDeviceSize size, const void* data)
{
    Cmd c{Op::UpdateBuffer};
    c.dst_buffer = reinterpret_cast<Buffer*>(dst);
    c.offset = offset;
    // the data is copied when the command is recorded (UpdateBufferCommand, CommandBuffer.cpp:241-276)
    c.bytes.assign((const uint8_t*)data, (const uint8_t*)data + size);
    CB(cb)->cmds.push_back(c);
}

void CmdClearColorImage(VkCommandBuffer cb, VkImage image, int32_t, const VkClearColorValue* color, uint32_t n,
                        const VkImageSubresourceRange* ranges)
{
    Cmd c{Op::ClearColorImage};
    c.dst_image = reinterpret_cast<Image*>(image);
    c.value.color = *color;
    c.ranges.assign(ranges, ranges + n);
    CB(cb)->cmds.push_back(c);
}

void CmdClearDepthStencilImage(VkCommandBuffer cb, VkImage image, int32_t, const VkClearDepthStencilValue* ds, uint32_t n,
                               const VkImageSubresourceRange* ranges)
{
    Cmd c{Op::ClearDepthStencilImage};
    c.dst_image = reinterpret_cast<Image*>(image);
    c.value.depthStencil = *ds;
    c.ranges.assign(ranges, ranges + n);
    CB(cb)->cmds.push_back(c);
}

void CmdClearAttachments(VkCommandBuffer cb, uint32_t n, const VkClearAttachment* atts, uint32_t nr, const VkClearRect* rects)
{
    Cmd c{Op::ClearAttachments};
    c.attachments.assign(atts, atts + n);
    c.rects.assign(rects, rects + nr);
    CB(cb)->cmds.push_back(c);
}

void CmdBlitImage(VkCommandBuffer cb, VkImage src, int32_t, VkImage dst, int32_t, uint32_t n, const VkImageBlit* regions, int32_t filter)
{
    Cmd c{Op::BlitImage};
    c.src_image = reinterpret_cast<Image*>(src);
    c.dst_image = reinterpret_cast<Image*>(dst);
    c.blits.assign(regions, regions + n);
    c.filter = filter;
    CB(cb)->cmds.push_back(c);
}

// ---------------------------------------------------------------- replay

struct ReplayState {
    RenderPass* pass = nullptr;
    Framebuffer* fb = nullptr;
    std::vector<VkClearValue> clears;
    Pipeline* pipeline = nullptr;
    Buffer* vb[CPV_MAX_VERTEX_BINDINGS] = {};
    VkDeviceSize vb_off[CPV_MAX_VERTEX_BINDINGS] = {};
    Buffer* ib = nullptr;
    VkDeviceSize ib_off = 0;
    int32_t ib_type = 0;
    DescriptorSet* set0 = nullptr;
    VkViewport viewport{};
    float blend_constants[4] = {};
    float depth_bounds[2] = {};
    uint32_t compare_mask[2] = {}, write_mask[2] = {}, reference[2] = {};   // front, back
    bool pass_split = false;   // a vkCmdClearAttachments ended an earlier part of the current pass
    std::vector<cpv_pipeline_desc> pipelines;
    std::vector<Pipeline*> pipeline_objs;
    std::vector<cpv_draw_desc> draws;
    std::vector<uint32_t> draw_pipeline;
};

// A render pass ready for libcpv, with its draws: the passes of a submit
// are gathered and handed over together (cpv_render_passes), so that
// independent passes share launches.
struct QueuedPass {
    cpv_pass_desc pass;
    std::vector<cpv_pipeline_desc> pipelines;
    std::vector<cpv_draw_desc> draws;
    std::vector<uint32_t> draw_pipeline;
};

// The flat arrays of cpv_render_passes / cpv_batch_create_passes.
struct PassArrays {
    std::vector<cpv_pass_desc> passes;
    std::vector<cpv_pipeline_desc> pipelines;
    std::vector<cpv_draw_desc> draws;
    std::vector<uint32_t> draw_pipeline, first_draw;
    explicit PassArrays(const std::vector<QueuedPass>& q)
    {
        first_draw.push_back(0);
        for (const QueuedPass& p : q) {
            const uint32_t base = (uint32_t)pipelines.size();
            passes.push_back(p.pass);
            pipelines.insert(pipelines.end(), p.pipelines.begin(), p.pipelines.end());
            draws.insert(draws.end(), p.draws.begin(), p.draws.end());
            for (uint32_t k : p.draw_pipeline) draw_pipeline.push_back(base + k);
            first_draw.push_back((uint32_t)draws.size());
        }
    }
};

VkResult flush_passes(Device* dev, std::vector<QueuedPass>& q)
{
    if (q.empty()) return VK_SUCCESS;
    const PassArrays a(q);
    q.clear();
    return device_error(cpv_render_passes(dev->ctx, a.passes.data(), (uint32_t)a.passes.size(), a.pipelines.data(), a.draws.data(),
                                          a.draw_pipeline.data(), a.first_draw.data()));
}

cpv_draw_desc make_draw(const ReplayState& s, const Cmd& c)
{
    cpv_draw_desc d;
    std::memset(&d, 0, sizeof(d));
    d.indexed = c.op == Op::DrawIndexed;
    d.count = c.count;
    d.instance_count = c.instances;
    d.first = c.first_vertex;
    d.vertex_offset = c.vertex_offset;
    d.first_instance = c.first_instance;
    if (d.indexed && s.ib) {
        d.index_size = s.ib_type == VK_INDEX_TYPE_UINT16 ? 2 : 4;
        d.index_address = s.ib->address() + s.ib_off;
    }
    for (int b = 0; b < CPV_MAX_VERTEX_BINDINGS; b++) {
        if (!s.vb[b]) continue;
        d.vertex_address[b] = s.vb[b]->address() + s.vb_off[b];
        d.vertex_bytes[b] = s.vb[b]->size > s.vb_off[b] ? s.vb[b]->size - s.vb_off[b] : 0;
    }
    const VkViewport& v = s.pipeline->dynamic_viewport ? s.viewport : s.pipeline->viewport;
    const float vp[6] = {v.x, v.y, v.width, v.height, v.minDepth, v.maxDepth};
    std::memcpy(d.viewport, vp, sizeof(vp));
    if (s.set0) {
        for (const auto& e : s.set0->bindings) {
            const uint32_t b = e.first;
            const Descriptor& ds = e.second;
            if (b >= CPV_MAX_UNIFORMS) continue;
            if (ds.type == VK_DESCRIPTOR_TYPE_UNIFORM_BUFFER && ds.buffer) d.uniform_address[b] = ds.buffer->address() + ds.offset;
            if (ds.type == VK_DESCRIPTOR_TYPE_COMBINED_IMAGE_SAMPLER && ds.view && ds.sampler) {
                cpv_texture_desc& t = d.textures[b];
                const Image* im = ds.view->image;
                t.address = im->address();
                t.format = (uint32_t)ds.view->format;
                t.width = im->width;
                t.height = im->height;
                t.levels = im->levels;
                t.base_level = ds.view->base_level;
                t.level_count = ds.view->level_count;
                t.sampler = ds.sampler->desc;
            }
        }
    }
    return d;
}

VkResult end_pass(ReplayState& s, std::vector<QueuedPass>* queue)
{
    const bool first_segment = !s.pass_split;
    cpv_pass_desc p;
    std::memset(&p, 0, sizeof(p));
    p.width = s.fb->width;
    p.height = s.fb->height;
    auto attach = [&](int32_t idx, cpv_attachment_desc* a) {
        const ImageView* v = s.fb->views.at((size_t)idx);
        const VkAttachmentDescription& ad = s.pass->attachments[(size_t)idx];
        a->address = v->image->address() + v->image->level_offset[v->base_level];
        a->format = (uint32_t)v->format;
        // a pass split by vkCmdClearAttachments continues with what its first part stored
        a->load_op = first_segment ? (uint32_t)ad.loadOp : (uint32_t)VK_ATTACHMENT_LOAD_OP_LOAD;
        if ((size_t)idx < s.clears.size()) std::memcpy(&a->clear, &s.clears[(size_t)idx], sizeof(VkClearValue));
    };
    if (s.pass->color >= 0) {
        p.color_count = 1;
        attach(s.pass->color, &p.color[0]);
    }
    if (s.pass->depth >= 0) {
        p.has_depth_stencil = 1;
        attach(s.pass->depth, &p.depth_stencil);
    }
    QueuedPass qp;
    qp.pass = p;
    qp.pipelines.swap(s.pipelines);
    qp.draws.swap(s.draws);
    qp.draw_pipeline.swap(s.draw_pipeline);
    s.pipeline_objs.clear();
    queue->push_back(std::move(qp));
    return VK_SUCCESS;
}

// ------------------------------------------------------ transfer replay

uint32_t level_width(const Image* im, uint32_t l) { return std::max(im->width >> l, 1u); }
uint32_t level_height(const Image* im, uint32_t l) { return std::max(im->height >> l, 1u); }

cpv_image_level image_level(const Image* im, uint32_t l)
{
    cpv_image_level v;
    std::memset(&v, 0, sizeof(v));
    v.address = im->address() + im->level_offset.at(l);
    v.format = (uint32_t)im->format;
    v.width = level_width(im, l);
    v.height = level_height(im, l);
    return v;
}

bool single_layer(const VkImageSubresourceLayers& r) { return r.baseArrayLayer == 0 && r.layerCount == 1; }

// Formats copied as plain pixel rows: colour formats and depth-only formats
// (ProcessNormal / the depth branch of ProcessDepthStencil,
// CommandBuffer.Copy.cpp:501-544, 546-620); combined depth/stencil formats
// take per-aspect repacking there and are not supported here.
bool row_copyable(int32_t f) { return !(f == CPV_FORMAT_D16_UNORM_S8_UINT || f == CPV_FORMAT_D24_UNORM_S8_UINT || f == CPV_FORMAT_D32_SFLOAT_S8_UINT || f == CPV_FORMAT_S8_UINT); }

// The rows of a VkBufferImageCopy region (CopyBufferToImageCommand::ProcessNormal,
// CommandBuffer.Copy.cpp:501-544, and CopyImageToBufferCommand::ProcessNormal :811-855):
// buffer row y of the region starts at bufferOffset + y * rowLength * pixel,
// image row y at level.Offset + (y + offset.y) * level.Stride + offset.x * pixel.
VkResult buffer_image_rows(const Buffer* buf, const Image* im, const VkBufferImageCopy& r, bool to_image, cpv_copy_region* out)
{
    std::memset(out, 0, sizeof(*out));   // zero rows: nothing to copy
    if (!single_layer(r.imageSubresource) || r.imageOffset.z != 0 || r.imageExtent.depth != 1 || !row_copyable(im->format) ||
        r.imageSubresource.mipLevel >= im->levels)
        return VK_ERROR_FEATURE_NOT_PRESENT;
    const uint64_t px = format_bytes(im->format);
    const uint32_t row_length = r.bufferRowLength ? r.bufferRowLength : r.imageExtent.width;
    const uint32_t l = r.imageSubresource.mipLevel;
    if (!r.imageExtent.width || !r.imageExtent.height) return VK_SUCCESS;
    if (r.imageOffset.x < 0 || r.imageOffset.y < 0 || (uint64_t)r.imageOffset.x + r.imageExtent.width > level_width(im, l) ||
        (uint64_t)r.imageOffset.y + r.imageExtent.height > level_height(im, l) || row_length < r.imageExtent.width ||
        r.bufferOffset + ((uint64_t)(r.imageExtent.height - 1) * row_length + r.imageExtent.width) * px > buf->size)
        return VK_ERROR_DEVICE_LOST;   // region outside the buffer or the image level
    const uint64_t stride = px * level_width(im, l);
    const uint64_t buf_at = buf->address() + r.bufferOffset;
    const uint64_t img_at = im->address() + im->level_offset[l] + (uint64_t)r.imageOffset.y * stride + (uint64_t)r.imageOffset.x * px;
    out->row_bytes = px * r.imageExtent.width;
    out->rows = r.imageExtent.height;
    out->src = to_image ? buf_at : img_at;
    out->dst = to_image ? img_at : buf_at;
    out->src_pitch = to_image ? px * row_length : stride;
    out->dst_pitch = to_image ? stride : px * row_length;
    return VK_SUCCESS;
}

// CopyImageCommand::Process for colour formats (CommandBuffer.Copy.cpp:77-196):
// rows of pixelSize * extent.width bytes at each level's Stride.
VkResult image_rows(const Image* src, const Image* dst, const VkImageCopy& r, cpv_copy_region* out)
{
    if (!single_layer(r.srcSubresource) || !single_layer(r.dstSubresource) || r.extent.depth != 1 || r.srcOffset.z || r.dstOffset.z ||
        !row_copyable(src->format) || format_bytes(src->format) != format_bytes(dst->format) ||
        r.srcSubresource.mipLevel >= src->levels || r.dstSubresource.mipLevel >= dst->levels)
        return VK_ERROR_FEATURE_NOT_PRESENT;
    const uint64_t px = format_bytes(src->format);
    const uint32_t sl = r.srcSubresource.mipLevel, dl = r.dstSubresource.mipLevel;
    if (r.srcOffset.x < 0 || r.srcOffset.y < 0 || r.dstOffset.x < 0 || r.dstOffset.y < 0 ||
        (uint64_t)r.srcOffset.x + r.extent.width > level_width(src, sl) || (uint64_t)r.srcOffset.y + r.extent.height > level_height(src, sl) ||
        (uint64_t)r.dstOffset.x + r.extent.width > level_width(dst, dl) || (uint64_t)r.dstOffset.y + r.extent.height > level_height(dst, dl))
        return VK_ERROR_DEVICE_LOST;   // region outside an image level
    std::memset(out, 0, sizeof(*out));
    out->src_pitch = px * level_width(src, sl);
    out->dst_pitch = px * level_width(dst, dl);
    out->src = src->address() + src->level_offset[sl] + (uint64_t)r.srcOffset.y * out->src_pitch + (uint64_t)r.srcOffset.x * px;
    out->dst = dst->address() + dst->level_offset[dl] + (uint64_t)r.dstOffset.y * out->dst_pitch + (uint64_t)r.dstOffset.x * px;
    out->row_bytes = px * r.extent.width;
    out->rows = r.extent.height;
    return VK_SUCCESS;
}

VkResult run_transfer(Device* dev, const Cmd& c)
{
    cpv_context* ctx = dev->ctx;
    std::vector<cpv_copy_region> rows;
    switch (c.op) {
    case Op::CopyBuffer:
        // CopyBufferCommand::Process (CommandBuffer.Copy.cpp:38-46): one memcpy per region
        for (const VkBufferCopy& r : c.buffer_copies) {
            if (r.srcOffset + r.size > c.src_buffer->size || r.dstOffset + r.size > c.dst_buffer->size) return VK_ERROR_DEVICE_LOST;
            cpv_copy_region g;
            std::memset(&g, 0, sizeof(g));
            g.src = c.src_buffer->address() + r.srcOffset;
            g.dst = c.dst_buffer->address() + r.dstOffset;
            g.src_pitch = g.dst_pitch = g.row_bytes = r.size;
            g.rows = 1;
            rows.push_back(g);
        }
        break;
    case Op::CopyImage:
        for (const VkImageCopy& r : c.image_copies) {
            cpv_copy_region g;
            if (VkResult e = image_rows(c.src_image, c.dst_image, r, &g)) return e;
            rows.push_back(g);
        }
        break;
    case Op::CopyBufferToImage: case Op::CopyImageToBuffer: {
        const bool to_image = c.op == Op::CopyBufferToImage;
        for (const VkBufferImageCopy& r : c.buffer_image_copies) {
            cpv_copy_region g;
            if (VkResult e = buffer_image_rows(to_image ? c.src_buffer : c.dst_buffer, to_image ? c.dst_image : c.src_image, r, to_image, &g))
                return e;
            rows.push_back(g);
        }
        break;
    }
    case Op::FillBuffer:
        if (c.offset + c.size > c.dst_buffer->size) return VK_ERROR_DEVICE_LOST;
        return device_error(cpv_fill_buffer(ctx, c.dst_buffer->address() + c.offset, c.size, c.data));
    case Op::UpdateBuffer:
        if (c.offset + c.bytes.size() > c.dst_buffer->size) return VK_ERROR_DEVICE_LOST;
        return device_error(cpv_update_buffer(ctx, c.dst_buffer->address() + c.offset, c.bytes.data(), c.bytes.size()));
    case Op::ClearColorImage: case Op::ClearDepthStencilImage: {
        // ClearImage over every level of each range (CommandBuffer.Draw.cpp:2398-2419, 117-149)
        uint32_t words[4];
        std::memcpy(words, &c.value, 16);
        for (const VkImageSubresourceRange& r : c.ranges) {
            const uint32_t levels = r.levelCount == VK_REMAINING_MIP_LEVELS ? c.dst_image->levels - r.baseMipLevel : r.levelCount;
            for (uint32_t l = r.baseMipLevel; l < r.baseMipLevel + levels && l < c.dst_image->levels; l++) {
                const cpv_image_level lv = image_level(c.dst_image, l);
                const cpv_rect rect{0, 0, lv.width, lv.height};
                if (VkResult e = device_error(cpv_clear_rects(ctx, &lv, words, &rect, 1))) return e;
            }
        }
        return VK_SUCCESS;
    }
    case Op::BlitImage: {
        for (const VkImageBlit& r : c.blits) {
            if (!single_layer(r.srcSubresource) || !single_layer(r.dstSubresource) || r.srcSubresource.mipLevel >= c.src_image->levels ||
                r.dstSubresource.mipLevel >= c.dst_image->levels)
                return VK_ERROR_FEATURE_NOT_PRESENT;
            const cpv_image_level sl = image_level(c.src_image, r.srcSubresource.mipLevel);
            const cpv_image_level dl = image_level(c.dst_image, r.dstSubresource.mipLevel);
            cpv_blit_region g;
            std::memset(&g, 0, sizeof(g));
            g.src[0] = r.srcOffsets[0].x; g.src[1] = r.srcOffsets[0].y; g.src[2] = r.srcOffsets[1].x; g.src[3] = r.srcOffsets[1].y;
            g.dst[0] = r.dstOffsets[0].x; g.dst[1] = r.dstOffsets[0].y; g.dst[2] = r.dstOffsets[1].x; g.dst[3] = r.dstOffsets[1].y;
            g.aspect = r.srcSubresource.aspectMask;
            if (VkResult e = device_error(cpv_blit_image(ctx, &sl, &dl, &g, 1, (uint32_t)c.filter))) return e;
        }
        return VK_SUCCESS;
    }
    default:
        return VK_ERROR_DEVICE_LOST;
    }
    return device_error(cpv_copy_regions(ctx, rows.data(), (uint32_t)rows.size()));
}

// ClearAttachmentsCommand::Process (CommandBuffer.Draw.cpp:2160-2224) inside a
// render pass: the draws recorded so far are issued first, then each
// rectangle of the attachment's first view level is set to the clear value
// (colour, or depth and stencil together, as the reference writes both).
VkResult clear_attachments(Device* dev, ReplayState& s, const Cmd& c, std::vector<QueuedPass>* queue)
{
    if (!s.pass || !s.fb) return VK_ERROR_DEVICE_LOST;
    if (VkResult e = end_pass(s, queue)) return e;
    if (VkResult e = flush_passes(dev, *queue)) return e;
    s.pass_split = true;
    for (const VkClearAttachment& a : c.attachments) {
        const int32_t idx = (a.aspectMask & VK_IMAGE_ASPECT_COLOR_BIT) ? s.pass->color : s.pass->depth;
        if (idx < 0) continue;
        const ImageView* v = s.fb->views.at((size_t)idx);
        const cpv_image_level lv = image_level(v->image, v->base_level);
        std::vector<cpv_rect> rects;
        for (const VkClearRect& r : c.rects) {
            if (r.baseArrayLayer != 0 || r.layerCount != 1) return VK_ERROR_FEATURE_NOT_PRESENT;
            rects.push_back(cpv_rect{r.rect.offset.x, r.rect.offset.y, r.rect.extent.width, r.rect.extent.height});
        }
        uint32_t words[4];
        std::memcpy(words, &a.clearValue, 16);
        if (VkResult e = device_error(cpv_clear_rects(dev->ctx, &lv, words, rects.data(), (uint32_t)rects.size()))) return e;
    }
    return VK_SUCCESS;
}

// DrawIndirectCommand / DrawIndexedIndirectCommand (CommandBuffer.Draw.cpp:1874-1996)
// read their VkDraw(Indexed)IndirectCommand records when the command runs;
// here they are read when the command buffer is replayed, after the stream
// has finished the work recorded before them, and become direct draws.
// Reads bytes of a buffer once the work enqueued before has finished.
VkResult read_buffer(Device* dev, const Buffer* b, VkDeviceSize offset, void* out, uint64_t bytes)
{
    if (offset + bytes > b->size) return VK_ERROR_DEVICE_LOST;
    if (b->memory && b->memory->host) {
        if (VkResult e = device_error(cpv_synchronize(dev->ctx))) return e;
        std::memcpy(out, (const void*)(uintptr_t)(b->address() + offset), bytes);
        return VK_SUCCESS;
    }
    return device_error(cpv_download(dev->ctx, out, b->address() + offset, bytes));
}

VkResult read_indirect(Device* dev, const Cmd& c, std::vector<Cmd>* out)
{
    const bool indexed = c.op == Op::DrawIndexedIndirect || c.op == Op::DrawIndexedIndirectCount;
    const bool counted = c.op == Op::DrawIndirectCount || c.op == Op::DrawIndexedIndirectCount;
    const uint32_t rec = indexed ? (uint32_t)sizeof(VkDrawIndexedIndirectCommand) : (uint32_t)sizeof(VkDrawIndirectCommand);
    uint32_t count = c.count;
    if (counted) {
        // min(maxDrawCount, the count buffer's value) records (CommandBuffer.Draw.cpp:2247, 2305)
        uint32_t n = 0;
        if (VkResult e = read_buffer(dev, c.count_buffer, c.count_offset, &n, 4)) return e;
        count = std::min(count, n);
    }
    const uint32_t stride = (counted || count > 1) && c.stride ? c.stride : rec;
    if (!count) return VK_SUCCESS;
    const uint64_t span = (uint64_t)stride * (count - 1) + rec;
    std::vector<uint8_t> host(span);
    if (VkResult e = read_buffer(dev, c.src_buffer, c.offset, host.data(), span)) return e;
    for (uint32_t j = 0; j < count; j++) {
        Cmd d{indexed ? Op::DrawIndexed : Op::Draw};
        if (indexed) {
            VkDrawIndexedIndirectCommand r;
            std::memcpy(&r, host.data() + (uint64_t)j * stride, sizeof(r));
            d.count = r.indexCount;
            d.instances = r.instanceCount;
            d.first_vertex = r.firstIndex;
            d.vertex_offset = r.vertexOffset;
            d.first_instance = r.firstInstance;
        } else {
            VkDrawIndirectCommand r;
            std::memcpy(&r, host.data() + (uint64_t)j * stride, sizeof(r));
            d.count = r.vertexCount;
            d.instances = r.instanceCount;
            d.first_vertex = r.firstVertex;
            d.first_instance = r.firstInstance;
        }
        out->push_back(d);
    }
    return VK_SUCCESS;
}

VkResult record_draw(ReplayState& s, const Cmd& c)
{
    if (!s.pass || !s.pipeline) return VK_ERROR_DEVICE_LOST;
    if (!c.count || !c.instances) return VK_SUCCESS;
    // the pipeline's state with its dynamic parts from the command buffer
    const Pipeline& p = *s.pipeline;
    cpv_pipeline_desc eff = p.desc;
    if (p.dyn_blend_constants) std::memcpy(eff.blend_constants, s.blend_constants, sizeof(eff.blend_constants));
    if (p.dyn_depth_bounds) {
        eff.min_depth_bounds = s.depth_bounds[0];
        eff.max_depth_bounds = s.depth_bounds[1];
    }
    cpv_stencil_state* face[2] = {&eff.front, &eff.back};
    for (int k = 0; k < 2; k++) {
        if (p.dyn_compare_mask) face[k]->compare_mask = s.compare_mask[k];
        if (p.dyn_write_mask) face[k]->write_mask = s.write_mask[k];
        if (p.dyn_reference) face[k]->reference = s.reference[k];
    }
    uint32_t pi = 0;
    for (; pi < s.pipelines.size(); pi++)
        if (s.pipeline_objs[pi] == s.pipeline && std::memcmp(&s.pipelines[pi], &eff, sizeof(eff)) == 0) break;
    if (pi == s.pipelines.size()) {
        s.pipeline_objs.push_back(s.pipeline);
        s.pipelines.push_back(eff);
    }
    s.draws.push_back(make_draw(s, c));
    s.draw_pipeline.push_back(pi);
    return VK_SUCCESS;
}

// Replays a command buffer: its render passes are queued for the submit's
// cpv_render_passes call; transfers, clears inside a pass and indirect
// reads first hand over the passes queued before them.
VkResult replay(Device* dev, const CommandBuffer* cb, std::vector<QueuedPass>* queue)
{
    ReplayState s;
    for (const Cmd& c : cb->cmds) {
        switch (c.op) {
        case Op::BeginPass:
            s.pass = c.pass;
            s.fb = c.fb;
            s.clears = c.clears;
            s.pass_split = false;
            break;
        case Op::EndPass: {
            const VkResult r = end_pass(s, queue);
            if (r != VK_SUCCESS) return r;
            s.pass = nullptr;
            break;
        }
        case Op::BindPipeline: s.pipeline = c.pipeline; break;
        case Op::BindVertex:
            for (size_t i = 0; i < c.buffers.size() && c.first + i < CPV_MAX_VERTEX_BINDINGS; i++) {
                s.vb[c.first + i] = c.buffers[i];
                s.vb_off[c.first + i] = c.offsets[i];
            }
            break;
        case Op::BindIndex:
            s.ib = c.buffers[0];
            s.ib_off = c.offsets[0];
            s.ib_type = c.index_type;
            break;
        case Op::BindSets:
            if (c.first == 0 && !c.sets.empty()) s.set0 = c.sets[0];
            break;
        case Op::SetViewport: s.viewport = c.viewport; break;
        case Op::SetBlendConstants: std::memcpy(s.blend_constants, c.values, sizeof(s.blend_constants)); break;
        case Op::SetDepthBounds: s.depth_bounds[0] = c.values[0]; s.depth_bounds[1] = c.values[1]; break;
        case Op::SetStencilCompareMask: case Op::SetStencilWriteMask: case Op::SetStencilReference: {
            uint32_t* v = c.op == Op::SetStencilCompareMask ? s.compare_mask : c.op == Op::SetStencilWriteMask ? s.write_mask : s.reference;
            if (c.first & VK_STENCIL_FACE_FRONT_BIT) v[0] = c.data;
            if (c.first & VK_STENCIL_FACE_BACK_BIT) v[1] = c.data;
            break;
        }
        case Op::Draw: case Op::DrawIndexed:
            if (!s.pass || !s.pipeline) return VK_ERROR_DEVICE_LOST;
            if (VkResult e = record_draw(s, c)) return e;
            break;
        case Op::DrawIndirect: case Op::DrawIndexedIndirect: case Op::DrawIndirectCount: case Op::DrawIndexedIndirectCount: {
            std::vector<Cmd> direct;
            if (VkResult e = read_indirect(dev, c, &direct)) return e;
            for (const Cmd& d : direct)
                if (VkResult e = record_draw(s, d)) return e;
            break;
        }
        case Op::ClearAttachments:
            if (VkResult e = clear_attachments(dev, s, c, queue)) return e;
            break;
        default:
            if (s.pass) return VK_ERROR_DEVICE_LOST;   // transfers are recorded outside render passes
            if (VkResult e = flush_passes(dev, *queue)) return e;
            if (VkResult e = run_transfer(dev, c)) return e;
            break;
        }
    }
    return VK_SUCCESS;
}

// A command buffer holding only render passes with direct draws and their
// state can be recorded once as a libcpv batch.
bool batchable(const CommandBuffer* cb)
{
    bool any_pass = false;
    for (const Cmd& c : cb->cmds) {
        switch (c.op) {
        case Op::BeginPass: any_pass = true; break;
        case Op::EndPass: case Op::BindPipeline: case Op::BindVertex: case Op::BindIndex: case Op::BindSets: case Op::SetViewport:
        case Op::Draw: case Op::DrawIndexed: case Op::SetBlendConstants: case Op::SetDepthBounds: case Op::SetStencilCompareMask:
        case Op::SetStencilWriteMask: case Op::SetStencilReference:
            break;
        default:
            return false;
        }
    }
    return any_pass;
}

// The batch of a command buffer recorded without ONE_TIME_SUBMIT: its
// render passes, recorded into hipGraphs on its first submit (with the
// context's tile configuration, scheduler and predictor at that time) and
// replayed by every later submit with no host work.
cpv_batch* command_buffer_batch(Device* dev, CommandBuffer* cb)
{
    if (cb->batch || cb->batch_tried || cb->one_time || !batchable(cb)) return cb->batch;
    cb->batch_tried = true;
    std::vector<QueuedPass> q;
    if (replay(dev, cb, &q) != VK_SUCCESS || q.empty()) return nullptr;
    const PassArrays a(q);
    cpv_batch* b = nullptr;
    if (cpv_batch_create_passes(dev->ctx, a.passes.data(), (uint32_t)a.passes.size(), a.pipelines.data(), a.draws.data(),
                                a.draw_pipeline.data(), a.first_draw.data(), &b) != CPV_OK)
        return nullptr;   // replayed pass by pass instead
    cb->batch = b;
    return b;
}

// Queue::Submit (CPVulkan/Queue.cpp:11-77) runs every command buffer and then
// signals the fence.  Here the render passes of all command buffers of the
// call are gathered and handed to libcpv together, a reusable command
// buffer replays its recorded batch, and the fence is a HIP event after the
// submit's work: nothing waits on the host.
VkResult QueueSubmit(VkQueue q, uint32_t n, const VkSubmitInfo* submits, VkFence fence)
{
    Device* dev = reinterpret_cast<Queue*>(q)->device;
    std::lock_guard<std::mutex> guard(dev->lock);
    uint64_t before = 0;
    const bool trace = std::getenv("CPVULKAN_AMD_TRACE_SUBMITS") != nullptr;
    if (trace) (void)cpv_launch_count(dev->ctx, &before);
    std::vector<QueuedPass> queue;
    for (uint32_t i = 0; i < n; i++) {
        for (uint32_t k = 0; k < submits[i].commandBufferCount; k++) {
            CommandBuffer* cb = reinterpret_cast<CommandBuffer*>(submits[i].pCommandBuffers[k]);
            if (cpv_batch* b = command_buffer_batch(dev, cb)) {
                if (VkResult e = flush_passes(dev, queue)) return e;
                if (VkResult e = device_error(cpv_batch_submit(b))) return e;
                continue;
            }
            if (VkResult e = replay(dev, cb, &queue)) return e;
        }
    }
    if (VkResult e = flush_passes(dev, queue)) return e;
    if (fence) {
        if (VkResult e = device_error(cpv_fence_signal(reinterpret_cast<Fence*>(fence)->f))) return e;
    }
    if (trace) {
        uint64_t after = 0;
        (void)cpv_launch_count(dev->ctx, &after);
        dev->submits++;
        std::fprintf(stderr, "cpvulkan_amd: submit %llu: %llu launches\n", (unsigned long long)dev->submits,
                     (unsigned long long)(after - before));
    }
    return VK_SUCCESS;
}

VkResult QueueWaitIdle(VkQueue q) { return device_error(cpv_synchronize(reinterpret_cast<Queue*>(q)->device->ctx)); }

VkResult CreateFence(VkDevice dev, const VkFenceCreateInfo* info, const VkAllocationCallbacks*, VkFence* out)
{
    Fence* f = new (std::nothrow) Fence();
    if (!f) return VK_ERROR_OUT_OF_HOST_MEMORY;
    const int signaled = (info->flags & VK_FENCE_CREATE_SIGNALED_BIT) != 0;
    if (cpv_fence_create(reinterpret_cast<Device*>(dev)->ctx, signaled, &f->f) != CPV_OK) {
        delete f;
        return VK_ERROR_OUT_OF_DEVICE_MEMORY;
    }
    *out = handle<VkFence>(f);
    return VK_SUCCESS;
}

void DestroyFence(VkDevice, VkFence fence, const VkAllocationCallbacks*)
{
    Fence* f = reinterpret_cast<Fence*>(fence);
    if (!f) return;
    cpv_fence_destroy(f->f);
    delete f;
}

VkResult fence_result(int rc)
{
    if (rc == CPV_NOT_READY) return VK_NOT_READY;
    if (rc == CPV_TIMEOUT) return VK_TIMEOUT;
    return device_error(rc);
}

VkResult GetFenceStatus(VkDevice, VkFence fence) { return fence_result(cpv_fence_status(reinterpret_cast<Fence*>(fence)->f)); }

// vkWaitForFences: all of the fences (waitAll) or any one, within timeout
// nanoseconds (VK_TIMEOUT after it; Device::WaitForFences,
// CPVulkan/Fence.cpp:192-200).
VkResult WaitForFences(VkDevice, uint32_t n, const VkFence* fences, VkBool32 waitAll, uint64_t timeout)
{
    if (n == 1) return fence_result(cpv_fence_wait(reinterpret_cast<Fence*>(fences[0])->f, timeout));
    const auto start = std::chrono::steady_clock::now();
    for (;;) {
        uint32_t done = 0;
        for (uint32_t i = 0; i < n; i++) {
            const int rc = cpv_fence_status(reinterpret_cast<Fence*>(fences[i])->f);
            if (rc < 0) return device_error(rc);
            if (rc == CPV_OK) done++;
        }
        if (waitAll ? done == n : done > 0) return VK_SUCCESS;
        const uint64_t waited =
            (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - start).count();
        if (waited >= timeout) return VK_TIMEOUT;
        std::this_thread::sleep_for(std::chrono::nanoseconds(std::min<uint64_t>(timeout - waited, 20000)));
    }
}

VkResult ResetFences(VkDevice, uint32_t n, const VkFence* fences)
{
    for (uint32_t i = 0; i < n; i++) {
        if (VkResult e = device_error(cpv_fence_reset(reinterpret_cast<Fence*>(fences[i])->f))) return e;
    }
    return VK_SUCCESS;
}

// --------------------------------------------------------------- dispatch

PFN_vkVoidFunction GetDeviceProcAddr(VkDevice, const char* name);

struct Entry {
    const char* name;
    PFN_vkVoidFunction fn;
};

#define E(n) {"vk" #n, reinterpret_cast<PFN_vkVoidFunction>(&n)}
const Entry kGlobal[] = {E(CreateInstance), {"vkEnumerateInstanceExtensionProperties", reinterpret_cast<PFN_vkVoidFunction>(&EnumerateExtensionProperties)},
                         {"vkEnumerateInstanceLayerProperties", reinterpret_cast<PFN_vkVoidFunction>(&EnumerateLayerProperties)}};
const Entry kInstance[] = {
    E(DestroyInstance), E(EnumeratePhysicalDevices), E(GetPhysicalDeviceProperties), E(GetPhysicalDeviceFeatures),
    E(GetPhysicalDeviceMemoryProperties), E(GetPhysicalDeviceQueueFamilyProperties), E(GetPhysicalDeviceFormatProperties),
    E(EnumerateDeviceExtensionProperties), E(CreateDevice), E(GetDeviceProcAddr),
};
const Entry kDevice[] = {
    E(DestroyDevice), E(GetDeviceQueue), E(DeviceWaitIdle), E(AllocateMemory), E(FreeMemory), E(MapMemory), E(UnmapMemory),
    E(FlushMappedMemoryRanges), E(InvalidateMappedMemoryRanges), E(CreateBuffer), E(DestroyBuffer), E(GetBufferMemoryRequirements),
    E(BindBufferMemory), E(CreateImage), E(DestroyImage), E(GetImageMemoryRequirements), E(BindImageMemory),
    E(GetImageSubresourceLayout), E(CreateImageView), E(DestroyImageView), E(CreateSampler), E(DestroySampler),
    E(CreateShaderModule), E(DestroyShaderModule), E(CreateDescriptorSetLayout), E(DestroyDescriptorSetLayout),
    E(CreatePipelineLayout), E(DestroyPipelineLayout), E(CreateDescriptorPool), E(DestroyDescriptorPool),
    E(AllocateDescriptorSets), E(FreeDescriptorSets), E(UpdateDescriptorSets), E(CreateRenderPass), E(DestroyRenderPass),
    E(CreateFramebuffer), E(DestroyFramebuffer), E(CreateGraphicsPipelines), E(DestroyPipeline), E(CreateCommandPool),
    E(DestroyCommandPool), E(AllocateCommandBuffers), E(FreeCommandBuffers), E(BeginCommandBuffer), E(EndCommandBuffer),
    E(ResetCommandBuffer), E(CmdBeginRenderPass), E(CmdEndRenderPass), E(CmdBindPipeline), E(CmdBindVertexBuffers),
    E(CmdBindIndexBuffer), E(CmdBindDescriptorSets), E(CmdSetViewport), E(CmdSetScissor), E(CmdSetDepthBounds),
    E(CmdSetLineWidth), E(CmdSetBlendConstants), E(CmdSetStencilCompareMask), E(CmdSetStencilWriteMask),
    E(CmdSetStencilReference), E(CmdDraw), E(CmdDrawIndexed),
    E(CmdDrawIndirect), E(CmdDrawIndexedIndirect),
    {"vkCmdDrawIndirectCountKHR", reinterpret_cast<PFN_vkVoidFunction>(&CmdDrawIndirectCount)},
    {"vkCmdDrawIndexedIndirectCountKHR", reinterpret_cast<PFN_vkVoidFunction>(&CmdDrawIndexedIndirectCount)},
    {"vkCmdDrawIndirectCountAMD", reinterpret_cast<PFN_vkVoidFunction>(&CmdDrawIndirectCount)},
    {"vkCmdDrawIndexedIndirectCountAMD", reinterpret_cast<PFN_vkVoidFunction>(&CmdDrawIndexedIndirectCount)},
    E(CmdCopyBuffer), E(CmdCopyImage), E(CmdCopyBufferToImage),
    E(CmdCopyImageToBuffer), E(CmdFillBuffer), E(CmdUpdateBuffer), E(CmdClearColorImage), E(CmdClearDepthStencilImage),
    E(CmdClearAttachments), E(CmdBlitImage),
    E(QueueSubmit), E(QueueWaitIdle), E(CreateFence), E(DestroyFence), E(WaitForFences), E(GetFenceStatus), E(ResetFences),
};
#undef E

template <size_t N> PFN_vkVoidFunction find(const Entry (&table)[N], const char* name)
{
    for (const Entry& e : table)
        if (std::strcmp(e.name, name) == 0) return e.fn;
    return nullptr;
}

PFN_vkVoidFunction GetDeviceProcAddr(VkDevice, const char* name)
{
    if (!name) return nullptr;
    return find(kDevice, name);
}

}  // namespace

extern "C" {

// Clamps the loader-ICD interface to version 5, as the reference does
// (CPVulkan/CPVulkan.cpp:95-104).
__attribute__((visibility("default"))) VkResult vk_icdNegotiateLoaderICDInterfaceVersion(uint32_t* pSupportedVersion)
{
    if (!pSupportedVersion) return VK_ERROR_INCOMPATIBLE_DRIVER;
    if (*pSupportedVersion > 5) *pSupportedVersion = 5;
    return VK_SUCCESS;
}

// Global, instance and device entry points by name (CPVulkan/CPVulkan.cpp:14-88).
__attribute__((visibility("default"))) PFN_vkVoidFunction vk_icdGetInstanceProcAddr(VkInstance instance, const char* pName)
{
    if (!pName) return nullptr;
    if (PFN_vkVoidFunction f = find(kGlobal, pName)) return f;
    if (std::strcmp(pName, "vkGetInstanceProcAddr") == 0) return reinterpret_cast<PFN_vkVoidFunction>(&vk_icdGetInstanceProcAddr);
    if (!instance) return nullptr;
    if (PFN_vkVoidFunction f = find(kInstance, pName)) return f;
    return find(kDevice, pName);
}

// No physical-device-level extensions (CPVulkan/CPVulkan.cpp:90-93).
__attribute__((visibility("default"))) PFN_vkVoidFunction vk_icdGetPhysicalDeviceProcAddr(VkInstance, const char*) { return nullptr; }

}
