// Vulkan ICD front end of the draw path: libcpvulkan_amd_icd.so.
//
// An application that points the loader at cpvulkan_amd_icd.json (or that
// dlopens this library and calls vk_icdGetInstanceProcAddr) gets a Vulkan 1.1
// device whose vkCmdDraw / vkCmdDrawIndexed, recorded into command buffers
// and replayed by vkQueueSubmit, run on the MI355X through the C-ABI of
// include/cpv.h.  The object model is the minimum the draw path needs; it
// mirrors the reference's (CPVulkan/*.cpp): commands are recorded into the
// command buffer and executed at submit (CommandBuffer.h:161,
// CommandBuffer.cpp:964-968); a render pass's draws are collected and, at
// vkCmdEndRenderPass, issued as one cpv_render_pass.
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
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
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
};

struct CommandPool {};

enum class Op { BeginPass, EndPass, BindPipeline, BindVertex, BindIndex, BindSets, SetViewport, Draw, DrawIndexed };

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
};

struct CommandBuffer : Dispatchable {
    Device* device = nullptr;
    std::vector<Cmd> cmds;
};

struct Fence {
    bool signaled = false;
    Device* device = nullptr;
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
    l.pointSizeRange[0] = l.pointSizeRange[1] = 1.0f;
    l.lineWidthRange[0] = l.lineWidthRange[1] = 1.0f;
    l.optimalBufferCopyOffsetAlignment = l.optimalBufferCopyRowPitchAlignment = l.nonCoherentAtomSize = 1;
}

void GetPhysicalDeviceFeatures(VkPhysicalDevice, VkPhysicalDeviceFeatures* f) { std::memset(f, 0, sizeof(*f)); }

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

VkResult EnumerateDeviceExtensionProperties(VkPhysicalDevice, const char*, uint32_t* count, VkExtensionProperties*)
{
    *count = 0;
    return VK_SUCCESS;
}

VkResult EnumerateLayerProperties(uint32_t* count, VkLayerProperties*)
{
    *count = 0;
    return VK_SUCCESS;
}

VkResult CreateDevice(VkPhysicalDevice, const VkDeviceCreateInfo* info, const VkAllocationCallbacks*, VkDevice* out)
{
    if (info && info->enabledExtensionCount) return VK_ERROR_EXTENSION_NOT_PRESENT;
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
    if (const VkPipelineRasterizationStateCreateInfo* r = ci.pRasterizationState) {
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
        for (uint32_t i = 0; i < ci.pDynamicState->dynamicStateCount; i++)
            if (ci.pDynamicState->pDynamicStates[i] == VK_DYNAMIC_STATE_VIEWPORT) p->dynamic_viewport = true;
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

VkResult BeginCommandBuffer(VkCommandBuffer cb, const VkCommandBufferBeginInfo*)
{
    reinterpret_cast<CommandBuffer*>(cb)->cmds.clear();
    return VK_SUCCESS;
}

VkResult EndCommandBuffer(VkCommandBuffer) { return VK_SUCCESS; }

VkResult ResetCommandBuffer(VkCommandBuffer cb, VkFlags)
{
    reinterpret_cast<CommandBuffer*>(cb)->cmds.clear();
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
    std::vector<cpv_pipeline_desc> pipelines;
    std::vector<Pipeline*> pipeline_objs;
    std::vector<cpv_draw_desc> draws;
    std::vector<uint32_t> draw_pipeline;
};

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

VkResult end_pass(Device* dev, ReplayState& s)
{
    cpv_pass_desc p;
    std::memset(&p, 0, sizeof(p));
    p.width = s.fb->width;
    p.height = s.fb->height;
    auto attach = [&](int32_t idx, cpv_attachment_desc* a) {
        const ImageView* v = s.fb->views.at((size_t)idx);
        const VkAttachmentDescription& ad = s.pass->attachments[(size_t)idx];
        a->address = v->image->address() + v->image->level_offset[v->base_level];
        a->format = (uint32_t)v->format;
        a->load_op = (uint32_t)ad.loadOp;
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
    const int rc = cpv_render_pass(dev->ctx, &p, s.pipelines.data(), s.draws.data(), s.draw_pipeline.data(), (uint32_t)s.draws.size());
    s.pipelines.clear();
    s.pipeline_objs.clear();
    s.draws.clear();
    s.draw_pipeline.clear();
    return device_error(rc);
}

VkResult replay(Device* dev, const CommandBuffer* cb)
{
    ReplayState s;
    for (const Cmd& c : cb->cmds) {
        switch (c.op) {
        case Op::BeginPass:
            s.pass = c.pass;
            s.fb = c.fb;
            s.clears = c.clears;
            break;
        case Op::EndPass: {
            const VkResult r = end_pass(dev, s);
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
        case Op::Draw: case Op::DrawIndexed: {
            if (!s.pass || !s.pipeline) return VK_ERROR_DEVICE_LOST;
            uint32_t pi = 0;
            for (; pi < s.pipeline_objs.size(); pi++)
                if (s.pipeline_objs[pi] == s.pipeline) break;
            if (pi == s.pipeline_objs.size()) {
                s.pipeline_objs.push_back(s.pipeline);
                s.pipelines.push_back(s.pipeline->desc);
            }
            s.draws.push_back(make_draw(s, c));
            s.draw_pipeline.push_back(pi);
            break;
        }
        }
    }
    return VK_SUCCESS;
}

VkResult QueueSubmit(VkQueue q, uint32_t n, const VkSubmitInfo* submits, VkFence fence)
{
    Device* dev = reinterpret_cast<Queue*>(q)->device;
    std::lock_guard<std::mutex> guard(dev->lock);
    for (uint32_t i = 0; i < n; i++) {
        for (uint32_t k = 0; k < submits[i].commandBufferCount; k++) {
            const VkResult r = replay(dev, reinterpret_cast<CommandBuffer*>(submits[i].pCommandBuffers[k]));
            if (r != VK_SUCCESS) return r;
        }
    }
    if (fence) {
        Fence* f = reinterpret_cast<Fence*>(fence);
        f->device = dev;
        f->signaled = false;   // signalled once the stream has drained (vkWaitForFences / vkGetFenceStatus)
    }
    return VK_SUCCESS;
}

VkResult QueueWaitIdle(VkQueue q) { return device_error(cpv_synchronize(reinterpret_cast<Queue*>(q)->device->ctx)); }

VkResult CreateFence(VkDevice, const VkFenceCreateInfo* info, const VkAllocationCallbacks*, VkFence* out)
{
    Fence* f = new Fence();
    f->signaled = (info->flags & VK_FENCE_CREATE_SIGNALED_BIT) != 0;
    *out = handle<VkFence>(f);
    return VK_SUCCESS;
}

void DestroyFence(VkDevice, VkFence f, const VkAllocationCallbacks*) { delete reinterpret_cast<Fence*>(f); }

VkResult WaitForFences(VkDevice, uint32_t n, const VkFence* fences, VkBool32, uint64_t)
{
    for (uint32_t i = 0; i < n; i++) {
        Fence* f = reinterpret_cast<Fence*>(fences[i]);
        if (!f->signaled && f->device) {
            const VkResult r = device_error(cpv_synchronize(f->device->ctx));
            if (r != VK_SUCCESS) return r;
        }
        f->signaled = true;
    }
    return VK_SUCCESS;
}

VkResult GetFenceStatus(VkDevice dev, VkFence fence)
{
    Fence* f = reinterpret_cast<Fence*>(fence);
    if (!f->signaled) return WaitForFences(dev, 1, &fence, VK_TRUE, ~0ull);
    return VK_SUCCESS;
}

VkResult ResetFences(VkDevice, uint32_t n, const VkFence* fences)
{
    for (uint32_t i = 0; i < n; i++) reinterpret_cast<Fence*>(fences[i])->signaled = false;
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
    E(CmdBindIndexBuffer), E(CmdBindDescriptorSets), E(CmdSetViewport), E(CmdSetScissor), E(CmdDraw), E(CmdDrawIndexed),
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
