/*
 * The subset of the Vulkan 1.1 API (64-bit ABI) that the draw-path ICD
 * implements: object creation for buffers, images, samplers, shader
 * modules, descriptor sets, render passes, framebuffers and graphics
 * pipelines; command buffers with vkCmdDraw / vkCmdDrawIndexed; queue
 * submission and fences.
 *
 * No Vulkan SDK headers exist in this image, so the declarations are the
 * project's own, written from the Vulkan 1.1 specification.  They are ABI:
 * an application compiled against the Khronos headers passes these structs
 * to the ICD, so every layout below is checked with a static size assertion
 * taken from the specification's member lists (64-bit: pointers and
 * VkDeviceSize are 8 bytes, handles are 64-bit).
 */
#ifndef CPV_VULKAN_H
#define CPV_VULKAN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#define CPV_VK_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define CPV_VK_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

#define VKAPI_ATTR
#define VKAPI_CALL
#define VKAPI_PTR

#define VK_MAKE_VERSION(major, minor, patch) (((major) << 22) | ((minor) << 12) | (patch))
#define VK_API_VERSION_1_1 VK_MAKE_VERSION(1, 1, 0)
#define VK_MAX_PHYSICAL_DEVICE_NAME_SIZE 256
#define VK_UUID_SIZE 16
#define VK_MAX_EXTENSION_NAME_SIZE 256
#define VK_MAX_DESCRIPTION_SIZE 256
#define VK_MAX_MEMORY_TYPES 32
#define VK_MAX_MEMORY_HEAPS 16
#define VK_WHOLE_SIZE (~0ULL)
#define VK_REMAINING_MIP_LEVELS (~0U)
#define VK_ATTACHMENT_UNUSED (~0U)
#define VK_TRUE 1
#define VK_FALSE 0
#define ICD_LOADER_MAGIC 0x01CDC0DE

typedef uint32_t VkFlags;
typedef uint32_t VkBool32;
typedef uint64_t VkDeviceSize;
typedef uint32_t VkSampleMask;

#define CPV_VK_HANDLE(name) typedef struct name##_T* name;
CPV_VK_HANDLE(VkInstance)
CPV_VK_HANDLE(VkPhysicalDevice)
CPV_VK_HANDLE(VkDevice)
CPV_VK_HANDLE(VkQueue)
CPV_VK_HANDLE(VkCommandBuffer)
/* non-dispatchable handles (64-bit on 64-bit platforms) */
CPV_VK_HANDLE(VkDeviceMemory)
CPV_VK_HANDLE(VkBuffer)
CPV_VK_HANDLE(VkImage)
CPV_VK_HANDLE(VkImageView)
CPV_VK_HANDLE(VkSampler)
CPV_VK_HANDLE(VkShaderModule)
CPV_VK_HANDLE(VkPipeline)
CPV_VK_HANDLE(VkPipelineLayout)
CPV_VK_HANDLE(VkPipelineCache)
CPV_VK_HANDLE(VkDescriptorSetLayout)
CPV_VK_HANDLE(VkDescriptorPool)
CPV_VK_HANDLE(VkDescriptorSet)
CPV_VK_HANDLE(VkRenderPass)
CPV_VK_HANDLE(VkFramebuffer)
CPV_VK_HANDLE(VkCommandPool)
CPV_VK_HANDLE(VkFence)
CPV_VK_HANDLE(VkSemaphore)
CPV_VK_HANDLE(VkBufferView)
#undef CPV_VK_HANDLE

typedef int32_t VkResult;
enum {
    VK_SUCCESS = 0, VK_NOT_READY = 1, VK_TIMEOUT = 2, VK_INCOMPLETE = 5,
    VK_ERROR_OUT_OF_HOST_MEMORY = -1, VK_ERROR_OUT_OF_DEVICE_MEMORY = -2, VK_ERROR_INITIALIZATION_FAILED = -3,
    VK_ERROR_DEVICE_LOST = -4, VK_ERROR_MEMORY_MAP_FAILED = -5, VK_ERROR_LAYER_NOT_PRESENT = -6,
    VK_ERROR_EXTENSION_NOT_PRESENT = -7, VK_ERROR_FEATURE_NOT_PRESENT = -8, VK_ERROR_INCOMPATIBLE_DRIVER = -9,
    VK_ERROR_TOO_MANY_OBJECTS = -10, VK_ERROR_FORMAT_NOT_SUPPORTED = -11
};

typedef int32_t VkStructureType;
enum {
    VK_STRUCTURE_TYPE_APPLICATION_INFO = 0, VK_STRUCTURE_TYPE_INSTANCE_CREATE_INFO = 1,
    VK_STRUCTURE_TYPE_DEVICE_QUEUE_CREATE_INFO = 2, VK_STRUCTURE_TYPE_DEVICE_CREATE_INFO = 3,
    VK_STRUCTURE_TYPE_SUBMIT_INFO = 4, VK_STRUCTURE_TYPE_MEMORY_ALLOCATE_INFO = 5,
    VK_STRUCTURE_TYPE_MAPPED_MEMORY_RANGE = 6, VK_STRUCTURE_TYPE_FENCE_CREATE_INFO = 8,
    VK_STRUCTURE_TYPE_BUFFER_CREATE_INFO = 12, VK_STRUCTURE_TYPE_IMAGE_CREATE_INFO = 14,
    VK_STRUCTURE_TYPE_IMAGE_VIEW_CREATE_INFO = 15, VK_STRUCTURE_TYPE_SHADER_MODULE_CREATE_INFO = 16,
    VK_STRUCTURE_TYPE_PIPELINE_SHADER_STAGE_CREATE_INFO = 18,
    VK_STRUCTURE_TYPE_PIPELINE_VERTEX_INPUT_STATE_CREATE_INFO = 19,
    VK_STRUCTURE_TYPE_PIPELINE_INPUT_ASSEMBLY_STATE_CREATE_INFO = 20,
    VK_STRUCTURE_TYPE_PIPELINE_VIEWPORT_STATE_CREATE_INFO = 22,
    VK_STRUCTURE_TYPE_PIPELINE_RASTERIZATION_STATE_CREATE_INFO = 23,
    VK_STRUCTURE_TYPE_PIPELINE_MULTISAMPLE_STATE_CREATE_INFO = 24,
    VK_STRUCTURE_TYPE_PIPELINE_DEPTH_STENCIL_STATE_CREATE_INFO = 25,
    VK_STRUCTURE_TYPE_PIPELINE_COLOR_BLEND_STATE_CREATE_INFO = 26,
    VK_STRUCTURE_TYPE_PIPELINE_DYNAMIC_STATE_CREATE_INFO = 27,
    VK_STRUCTURE_TYPE_GRAPHICS_PIPELINE_CREATE_INFO = 28, VK_STRUCTURE_TYPE_PIPELINE_LAYOUT_CREATE_INFO = 30,
    VK_STRUCTURE_TYPE_SAMPLER_CREATE_INFO = 31, VK_STRUCTURE_TYPE_DESCRIPTOR_SET_LAYOUT_CREATE_INFO = 32,
    VK_STRUCTURE_TYPE_DESCRIPTOR_POOL_CREATE_INFO = 33, VK_STRUCTURE_TYPE_DESCRIPTOR_SET_ALLOCATE_INFO = 34,
    VK_STRUCTURE_TYPE_WRITE_DESCRIPTOR_SET = 35, VK_STRUCTURE_TYPE_FRAMEBUFFER_CREATE_INFO = 37,
    VK_STRUCTURE_TYPE_RENDER_PASS_CREATE_INFO = 38, VK_STRUCTURE_TYPE_COMMAND_POOL_CREATE_INFO = 39,
    VK_STRUCTURE_TYPE_COMMAND_BUFFER_ALLOCATE_INFO = 40, VK_STRUCTURE_TYPE_COMMAND_BUFFER_BEGIN_INFO = 42,
    VK_STRUCTURE_TYPE_RENDER_PASS_BEGIN_INFO = 43
};

enum { VK_PHYSICAL_DEVICE_TYPE_DISCRETE_GPU = 2 };
enum { VK_QUEUE_GRAPHICS_BIT = 1, VK_QUEUE_COMPUTE_BIT = 2, VK_QUEUE_TRANSFER_BIT = 4 };
enum { VK_MEMORY_PROPERTY_DEVICE_LOCAL_BIT = 1, VK_MEMORY_PROPERTY_HOST_VISIBLE_BIT = 2,
       VK_MEMORY_PROPERTY_HOST_COHERENT_BIT = 4, VK_MEMORY_PROPERTY_HOST_CACHED_BIT = 8 };
enum { VK_MEMORY_HEAP_DEVICE_LOCAL_BIT = 1 };
enum { VK_IMAGE_TILING_OPTIMAL = 0, VK_IMAGE_TILING_LINEAR = 1 };
enum { VK_IMAGE_TYPE_2D = 1 };
enum { VK_DESCRIPTOR_TYPE_COMBINED_IMAGE_SAMPLER = 1, VK_DESCRIPTOR_TYPE_UNIFORM_BUFFER = 6 };
enum { VK_DYNAMIC_STATE_VIEWPORT = 0, VK_DYNAMIC_STATE_SCISSOR = 1, VK_DYNAMIC_STATE_LINE_WIDTH = 2, VK_DYNAMIC_STATE_DEPTH_BIAS = 3,
       VK_DYNAMIC_STATE_BLEND_CONSTANTS = 4, VK_DYNAMIC_STATE_DEPTH_BOUNDS = 5, VK_DYNAMIC_STATE_STENCIL_COMPARE_MASK = 6,
       VK_DYNAMIC_STATE_STENCIL_WRITE_MASK = 7, VK_DYNAMIC_STATE_STENCIL_REFERENCE = 8 };
enum { VK_STENCIL_FACE_FRONT_BIT = 1, VK_STENCIL_FACE_BACK_BIT = 2, VK_STENCIL_FACE_FRONT_AND_BACK = 3 };
enum { VK_COMMAND_BUFFER_USAGE_ONE_TIME_SUBMIT_BIT = 1 };
enum { VK_INDEX_TYPE_UINT16 = 0, VK_INDEX_TYPE_UINT32 = 1 };
enum { VK_IMAGE_ASPECT_COLOR_BIT = 1, VK_IMAGE_ASPECT_DEPTH_BIT = 2, VK_IMAGE_ASPECT_STENCIL_BIT = 4 };
enum { VK_FILTER_NEAREST = 0, VK_FILTER_LINEAR = 1 };
enum { VK_SHADER_STAGE_VERTEX_BIT = 1, VK_SHADER_STAGE_FRAGMENT_BIT = 0x10 };
enum { VK_PIPELINE_BIND_POINT_GRAPHICS = 0 };
enum { VK_ATTACHMENT_LOAD_OP_LOAD = 0, VK_ATTACHMENT_LOAD_OP_CLEAR = 1, VK_ATTACHMENT_LOAD_OP_DONT_CARE = 2 };
enum { VK_SUBPASS_CONTENTS_INLINE = 0 };
enum { VK_FENCE_CREATE_SIGNALED_BIT = 1 };
enum { VK_COMMAND_BUFFER_LEVEL_PRIMARY = 0 };

typedef struct VkAllocationCallbacks VkAllocationCallbacks;   /* accepted and ignored */
typedef void (VKAPI_PTR* PFN_vkVoidFunction)(void);

typedef struct VkExtent2D { uint32_t width, height; } VkExtent2D;
typedef struct VkExtent3D { uint32_t width, height, depth; } VkExtent3D;
typedef struct VkOffset2D { int32_t x, y; } VkOffset2D;
typedef struct VkRect2D { VkOffset2D offset; VkExtent2D extent; } VkRect2D;
typedef struct VkViewport { float x, y, width, height, minDepth, maxDepth; } VkViewport;

typedef struct VkApplicationInfo {
    VkStructureType sType; const void* pNext; const char* pApplicationName; uint32_t applicationVersion;
    const char* pEngineName; uint32_t engineVersion; uint32_t apiVersion;
} VkApplicationInfo;
typedef struct VkInstanceCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; const VkApplicationInfo* pApplicationInfo;
    uint32_t enabledLayerCount; const char* const* ppEnabledLayerNames;
    uint32_t enabledExtensionCount; const char* const* ppEnabledExtensionNames;
} VkInstanceCreateInfo;

typedef struct VkPhysicalDeviceLimits {
    uint32_t maxImageDimension1D, maxImageDimension2D, maxImageDimension3D, maxImageDimensionCube, maxImageArrayLayers;
    uint32_t maxTexelBufferElements, maxUniformBufferRange, maxStorageBufferRange, maxPushConstantsSize;
    uint32_t maxMemoryAllocationCount, maxSamplerAllocationCount;
    VkDeviceSize bufferImageGranularity, sparseAddressSpaceSize;
    uint32_t maxBoundDescriptorSets, maxPerStageDescriptorSamplers, maxPerStageDescriptorUniformBuffers;
    uint32_t maxPerStageDescriptorStorageBuffers, maxPerStageDescriptorSampledImages, maxPerStageDescriptorStorageImages;
    uint32_t maxPerStageDescriptorInputAttachments, maxPerStageResources, maxDescriptorSetSamplers;
    uint32_t maxDescriptorSetUniformBuffers, maxDescriptorSetUniformBuffersDynamic, maxDescriptorSetStorageBuffers;
    uint32_t maxDescriptorSetStorageBuffersDynamic, maxDescriptorSetSampledImages, maxDescriptorSetStorageImages;
    uint32_t maxDescriptorSetInputAttachments, maxVertexInputAttributes, maxVertexInputBindings;
    uint32_t maxVertexInputAttributeOffset, maxVertexInputBindingStride, maxVertexOutputComponents;
    uint32_t maxTessellationGenerationLevel, maxTessellationPatchSize, maxTessellationControlPerVertexInputComponents;
    uint32_t maxTessellationControlPerVertexOutputComponents, maxTessellationControlPerPatchOutputComponents;
    uint32_t maxTessellationControlTotalOutputComponents, maxTessellationEvaluationInputComponents;
    uint32_t maxTessellationEvaluationOutputComponents, maxGeometryShaderInvocations, maxGeometryInputComponents;
    uint32_t maxGeometryOutputComponents, maxGeometryOutputVertices, maxGeometryTotalOutputComponents;
    uint32_t maxFragmentInputComponents, maxFragmentOutputAttachments, maxFragmentDualSrcAttachments;
    uint32_t maxFragmentCombinedOutputResources, maxComputeSharedMemorySize, maxComputeWorkGroupCount[3];
    uint32_t maxComputeWorkGroupInvocations, maxComputeWorkGroupSize[3], subPixelPrecisionBits, subTexelPrecisionBits;
    uint32_t mipmapPrecisionBits, maxDrawIndexedIndexValue, maxDrawIndirectCount;
    float maxSamplerLodBias, maxSamplerAnisotropy;
    uint32_t maxViewports, maxViewportDimensions[2];
    float viewportBoundsRange[2];
    uint32_t viewportSubPixelBits;
    size_t minMemoryMapAlignment;
    VkDeviceSize minTexelBufferOffsetAlignment, minUniformBufferOffsetAlignment, minStorageBufferOffsetAlignment;
    int32_t minTexelOffset; uint32_t maxTexelOffset; int32_t minTexelGatherOffset; uint32_t maxTexelGatherOffset;
    float minInterpolationOffset, maxInterpolationOffset;
    uint32_t subPixelInterpolationOffsetBits, maxFramebufferWidth, maxFramebufferHeight, maxFramebufferLayers;
    VkFlags framebufferColorSampleCounts, framebufferDepthSampleCounts, framebufferStencilSampleCounts;
    VkFlags framebufferNoAttachmentsSampleCounts;
    uint32_t maxColorAttachments;
    VkFlags sampledImageColorSampleCounts, sampledImageIntegerSampleCounts, sampledImageDepthSampleCounts;
    VkFlags sampledImageStencilSampleCounts, storageImageSampleCounts;
    uint32_t maxSampleMaskWords;
    VkBool32 timestampComputeAndGraphics;
    float timestampPeriod;
    uint32_t maxClipDistances, maxCullDistances, maxCombinedClipAndCullDistances, discreteQueuePriorities;
    float pointSizeRange[2], lineWidthRange[2], pointSizeGranularity, lineWidthGranularity;
    VkBool32 strictLines, standardSampleLocations;
    VkDeviceSize optimalBufferCopyOffsetAlignment, optimalBufferCopyRowPitchAlignment, nonCoherentAtomSize;
} VkPhysicalDeviceLimits;
typedef struct VkPhysicalDeviceSparseProperties {
    VkBool32 residencyStandard2DBlockShape, residencyStandard2DMultisampleBlockShape, residencyStandard3DBlockShape;
    VkBool32 residencyAlignedMipSize, residencyNonResidentStrict;
} VkPhysicalDeviceSparseProperties;
typedef struct VkPhysicalDeviceProperties {
    uint32_t apiVersion, driverVersion, vendorID, deviceID;
    int32_t deviceType;
    char deviceName[VK_MAX_PHYSICAL_DEVICE_NAME_SIZE];
    uint8_t pipelineCacheUUID[VK_UUID_SIZE];
    VkPhysicalDeviceLimits limits;
    VkPhysicalDeviceSparseProperties sparseProperties;
} VkPhysicalDeviceProperties;
typedef struct VkPhysicalDeviceFeatures { VkBool32 feature[55]; } VkPhysicalDeviceFeatures;
typedef struct VkMemoryType { VkFlags propertyFlags; uint32_t heapIndex; } VkMemoryType;
typedef struct VkMemoryHeap { VkDeviceSize size; VkFlags flags; } VkMemoryHeap;
typedef struct VkPhysicalDeviceMemoryProperties {
    uint32_t memoryTypeCount; VkMemoryType memoryTypes[VK_MAX_MEMORY_TYPES];
    uint32_t memoryHeapCount; VkMemoryHeap memoryHeaps[VK_MAX_MEMORY_HEAPS];
} VkPhysicalDeviceMemoryProperties;
typedef struct VkQueueFamilyProperties {
    VkFlags queueFlags; uint32_t queueCount; uint32_t timestampValidBits; VkExtent3D minImageTransferGranularity;
} VkQueueFamilyProperties;
typedef struct VkExtensionProperties { char extensionName[VK_MAX_EXTENSION_NAME_SIZE]; uint32_t specVersion; } VkExtensionProperties;
typedef struct VkLayerProperties {
    char layerName[VK_MAX_EXTENSION_NAME_SIZE]; uint32_t specVersion; uint32_t implementationVersion;
    char description[VK_MAX_DESCRIPTION_SIZE];
} VkLayerProperties;
typedef struct VkFormatProperties { VkFlags linearTilingFeatures, optimalTilingFeatures, bufferFeatures; } VkFormatProperties;

typedef struct VkDeviceQueueCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; uint32_t queueFamilyIndex; uint32_t queueCount;
    const float* pQueuePriorities;
} VkDeviceQueueCreateInfo;
typedef struct VkDeviceCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; uint32_t queueCreateInfoCount;
    const VkDeviceQueueCreateInfo* pQueueCreateInfos; uint32_t enabledLayerCount; const char* const* ppEnabledLayerNames;
    uint32_t enabledExtensionCount; const char* const* ppEnabledExtensionNames; const VkPhysicalDeviceFeatures* pEnabledFeatures;
} VkDeviceCreateInfo;

typedef struct VkMemoryAllocateInfo {
    VkStructureType sType; const void* pNext; VkDeviceSize allocationSize; uint32_t memoryTypeIndex;
} VkMemoryAllocateInfo;
typedef struct VkMemoryRequirements { VkDeviceSize size, alignment; uint32_t memoryTypeBits; } VkMemoryRequirements;
typedef struct VkMappedMemoryRange {
    VkStructureType sType; const void* pNext; VkDeviceMemory memory; VkDeviceSize offset, size;
} VkMappedMemoryRange;

typedef struct VkBufferCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; VkDeviceSize size; VkFlags usage; int32_t sharingMode;
    uint32_t queueFamilyIndexCount; const uint32_t* pQueueFamilyIndices;
} VkBufferCreateInfo;
typedef struct VkImageCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; int32_t imageType; int32_t format; VkExtent3D extent;
    uint32_t mipLevels, arrayLayers; VkFlags samples; int32_t tiling; VkFlags usage; int32_t sharingMode;
    uint32_t queueFamilyIndexCount; const uint32_t* pQueueFamilyIndices; int32_t initialLayout;
} VkImageCreateInfo;
typedef struct VkImageSubresource { VkFlags aspectMask; uint32_t mipLevel, arrayLayer; } VkImageSubresource;
typedef struct VkSubresourceLayout { VkDeviceSize offset, size, rowPitch, arrayPitch, depthPitch; } VkSubresourceLayout;
typedef struct VkComponentMapping { int32_t r, g, b, a; } VkComponentMapping;
typedef struct VkImageSubresourceRange {
    VkFlags aspectMask; uint32_t baseMipLevel, levelCount, baseArrayLayer, layerCount;
} VkImageSubresourceRange;
typedef struct VkImageViewCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; VkImage image; int32_t viewType; int32_t format;
    VkComponentMapping components; VkImageSubresourceRange subresourceRange;
} VkImageViewCreateInfo;
typedef struct VkSamplerCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; int32_t magFilter, minFilter, mipmapMode;
    int32_t addressModeU, addressModeV, addressModeW; float mipLodBias; VkBool32 anisotropyEnable; float maxAnisotropy;
    VkBool32 compareEnable; int32_t compareOp; float minLod, maxLod; int32_t borderColor; VkBool32 unnormalizedCoordinates;
} VkSamplerCreateInfo;
typedef struct VkShaderModuleCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; size_t codeSize; const uint32_t* pCode;
} VkShaderModuleCreateInfo;

typedef struct VkDescriptorSetLayoutBinding {
    uint32_t binding; int32_t descriptorType; uint32_t descriptorCount; VkFlags stageFlags; const VkSampler* pImmutableSamplers;
} VkDescriptorSetLayoutBinding;
typedef struct VkDescriptorSetLayoutCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; uint32_t bindingCount; const VkDescriptorSetLayoutBinding* pBindings;
} VkDescriptorSetLayoutCreateInfo;
typedef struct VkPushConstantRange { VkFlags stageFlags; uint32_t offset, size; } VkPushConstantRange;
typedef struct VkPipelineLayoutCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; uint32_t setLayoutCount; const VkDescriptorSetLayout* pSetLayouts;
    uint32_t pushConstantRangeCount; const VkPushConstantRange* pPushConstantRanges;
} VkPipelineLayoutCreateInfo;
typedef struct VkDescriptorPoolSize { int32_t type; uint32_t descriptorCount; } VkDescriptorPoolSize;
typedef struct VkDescriptorPoolCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; uint32_t maxSets; uint32_t poolSizeCount;
    const VkDescriptorPoolSize* pPoolSizes;
} VkDescriptorPoolCreateInfo;
typedef struct VkDescriptorSetAllocateInfo {
    VkStructureType sType; const void* pNext; VkDescriptorPool descriptorPool; uint32_t descriptorSetCount;
    const VkDescriptorSetLayout* pSetLayouts;
} VkDescriptorSetAllocateInfo;
typedef struct VkDescriptorImageInfo { VkSampler sampler; VkImageView imageView; int32_t imageLayout; } VkDescriptorImageInfo;
typedef struct VkDescriptorBufferInfo { VkBuffer buffer; VkDeviceSize offset, range; } VkDescriptorBufferInfo;
typedef struct VkWriteDescriptorSet {
    VkStructureType sType; const void* pNext; VkDescriptorSet dstSet; uint32_t dstBinding, dstArrayElement, descriptorCount;
    int32_t descriptorType; const VkDescriptorImageInfo* pImageInfo; const VkDescriptorBufferInfo* pBufferInfo;
    const VkBufferView* pTexelBufferView;
} VkWriteDescriptorSet;
typedef struct VkCopyDescriptorSet VkCopyDescriptorSet;

typedef struct VkAttachmentDescription {
    VkFlags flags; int32_t format; VkFlags samples; int32_t loadOp, storeOp, stencilLoadOp, stencilStoreOp;
    int32_t initialLayout, finalLayout;
} VkAttachmentDescription;
typedef struct VkAttachmentReference { uint32_t attachment; int32_t layout; } VkAttachmentReference;
typedef struct VkSubpassDescription {
    VkFlags flags; int32_t pipelineBindPoint; uint32_t inputAttachmentCount; const VkAttachmentReference* pInputAttachments;
    uint32_t colorAttachmentCount; const VkAttachmentReference* pColorAttachments; const VkAttachmentReference* pResolveAttachments;
    const VkAttachmentReference* pDepthStencilAttachment; uint32_t preserveAttachmentCount; const uint32_t* pPreserveAttachments;
} VkSubpassDescription;
typedef struct VkSubpassDependency {
    uint32_t srcSubpass, dstSubpass; VkFlags srcStageMask, dstStageMask, srcAccessMask, dstAccessMask, dependencyFlags;
} VkSubpassDependency;
typedef struct VkRenderPassCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; uint32_t attachmentCount; const VkAttachmentDescription* pAttachments;
    uint32_t subpassCount; const VkSubpassDescription* pSubpasses; uint32_t dependencyCount; const VkSubpassDependency* pDependencies;
} VkRenderPassCreateInfo;
typedef struct VkFramebufferCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; VkRenderPass renderPass; uint32_t attachmentCount;
    const VkImageView* pAttachments; uint32_t width, height, layers;
} VkFramebufferCreateInfo;

typedef struct VkSpecializationInfo VkSpecializationInfo;
typedef struct VkPipelineShaderStageCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; VkFlags stage; VkShaderModule module; const char* pName;
    const VkSpecializationInfo* pSpecializationInfo;
} VkPipelineShaderStageCreateInfo;
typedef struct VkVertexInputBindingDescription { uint32_t binding, stride; int32_t inputRate; } VkVertexInputBindingDescription;
typedef struct VkVertexInputAttributeDescription { uint32_t location, binding; int32_t format; uint32_t offset; } VkVertexInputAttributeDescription;
typedef struct VkPipelineVertexInputStateCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; uint32_t vertexBindingDescriptionCount;
    const VkVertexInputBindingDescription* pVertexBindingDescriptions; uint32_t vertexAttributeDescriptionCount;
    const VkVertexInputAttributeDescription* pVertexAttributeDescriptions;
} VkPipelineVertexInputStateCreateInfo;
typedef struct VkPipelineInputAssemblyStateCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; int32_t topology; VkBool32 primitiveRestartEnable;
} VkPipelineInputAssemblyStateCreateInfo;
typedef struct VkPipelineTessellationStateCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; uint32_t patchControlPoints;
} VkPipelineTessellationStateCreateInfo;
typedef struct VkPipelineViewportStateCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; uint32_t viewportCount; const VkViewport* pViewports;
    uint32_t scissorCount; const VkRect2D* pScissors;
} VkPipelineViewportStateCreateInfo;
typedef struct VkPipelineRasterizationStateCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; VkBool32 depthClampEnable, rasterizerDiscardEnable;
    int32_t polygonMode; VkFlags cullMode; int32_t frontFace; VkBool32 depthBiasEnable; float depthBiasConstantFactor;
    float depthBiasClamp, depthBiasSlopeFactor, lineWidth;
} VkPipelineRasterizationStateCreateInfo;
typedef struct VkPipelineMultisampleStateCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; VkFlags rasterizationSamples; VkBool32 sampleShadingEnable;
    float minSampleShading; const VkSampleMask* pSampleMask; VkBool32 alphaToCoverageEnable, alphaToOneEnable;
} VkPipelineMultisampleStateCreateInfo;
typedef struct VkStencilOpState {
    int32_t failOp, passOp, depthFailOp, compareOp; uint32_t compareMask, writeMask, reference;
} VkStencilOpState;
typedef struct VkPipelineDepthStencilStateCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; VkBool32 depthTestEnable, depthWriteEnable; int32_t depthCompareOp;
    VkBool32 depthBoundsTestEnable, stencilTestEnable; VkStencilOpState front, back; float minDepthBounds, maxDepthBounds;
} VkPipelineDepthStencilStateCreateInfo;
typedef struct VkPipelineColorBlendAttachmentState {
    VkBool32 blendEnable; int32_t srcColorBlendFactor, dstColorBlendFactor, colorBlendOp;
    int32_t srcAlphaBlendFactor, dstAlphaBlendFactor, alphaBlendOp; VkFlags colorWriteMask;
} VkPipelineColorBlendAttachmentState;
typedef struct VkPipelineColorBlendStateCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; VkBool32 logicOpEnable; int32_t logicOp; uint32_t attachmentCount;
    const VkPipelineColorBlendAttachmentState* pAttachments; float blendConstants[4];
} VkPipelineColorBlendStateCreateInfo;
typedef struct VkPipelineDynamicStateCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; uint32_t dynamicStateCount; const int32_t* pDynamicStates;
} VkPipelineDynamicStateCreateInfo;
typedef struct VkGraphicsPipelineCreateInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; uint32_t stageCount; const VkPipelineShaderStageCreateInfo* pStages;
    const VkPipelineVertexInputStateCreateInfo* pVertexInputState;
    const VkPipelineInputAssemblyStateCreateInfo* pInputAssemblyState;
    const VkPipelineTessellationStateCreateInfo* pTessellationState;
    const VkPipelineViewportStateCreateInfo* pViewportState;
    const VkPipelineRasterizationStateCreateInfo* pRasterizationState;
    const VkPipelineMultisampleStateCreateInfo* pMultisampleState;
    const VkPipelineDepthStencilStateCreateInfo* pDepthStencilState;
    const VkPipelineColorBlendStateCreateInfo* pColorBlendState;
    const VkPipelineDynamicStateCreateInfo* pDynamicState;
    VkPipelineLayout layout; VkRenderPass renderPass; uint32_t subpass; VkPipeline basePipelineHandle; int32_t basePipelineIndex;
} VkGraphicsPipelineCreateInfo;

typedef struct VkCommandPoolCreateInfo { VkStructureType sType; const void* pNext; VkFlags flags; uint32_t queueFamilyIndex; } VkCommandPoolCreateInfo;
typedef struct VkCommandBufferAllocateInfo {
    VkStructureType sType; const void* pNext; VkCommandPool commandPool; int32_t level; uint32_t commandBufferCount;
} VkCommandBufferAllocateInfo;
typedef struct VkCommandBufferInheritanceInfo VkCommandBufferInheritanceInfo;
typedef struct VkCommandBufferBeginInfo {
    VkStructureType sType; const void* pNext; VkFlags flags; const VkCommandBufferInheritanceInfo* pInheritanceInfo;
} VkCommandBufferBeginInfo;
typedef union VkClearColorValue { float float32[4]; int32_t int32[4]; uint32_t uint32[4]; } VkClearColorValue;
typedef struct VkClearDepthStencilValue { float depth; uint32_t stencil; } VkClearDepthStencilValue;
typedef union VkClearValue { VkClearColorValue color; VkClearDepthStencilValue depthStencil; } VkClearValue;
typedef struct VkRenderPassBeginInfo {
    VkStructureType sType; const void* pNext; VkRenderPass renderPass; VkFramebuffer framebuffer; VkRect2D renderArea;
    uint32_t clearValueCount; const VkClearValue* pClearValues;
} VkRenderPassBeginInfo;
typedef struct VkSubmitInfo {
    VkStructureType sType; const void* pNext; uint32_t waitSemaphoreCount; const VkSemaphore* pWaitSemaphores;
    const VkFlags* pWaitDstStageMask; uint32_t commandBufferCount; const VkCommandBuffer* pCommandBuffers;
    uint32_t signalSemaphoreCount; const VkSemaphore* pSignalSemaphores;
} VkSubmitInfo;
typedef struct VkFenceCreateInfo { VkStructureType sType; const void* pNext; VkFlags flags; } VkFenceCreateInfo;

/* Transfer, clear and indirect-draw commands. */
typedef struct VkOffset3D { int32_t x, y, z; } VkOffset3D;
typedef struct VkImageSubresourceLayers { VkFlags aspectMask; uint32_t mipLevel, baseArrayLayer, layerCount; } VkImageSubresourceLayers;
typedef struct VkBufferCopy { VkDeviceSize srcOffset, dstOffset, size; } VkBufferCopy;
typedef struct VkImageCopy {
    VkImageSubresourceLayers srcSubresource; VkOffset3D srcOffset; VkImageSubresourceLayers dstSubresource; VkOffset3D dstOffset;
    VkExtent3D extent;
} VkImageCopy;
typedef struct VkBufferImageCopy {
    VkDeviceSize bufferOffset; uint32_t bufferRowLength, bufferImageHeight; VkImageSubresourceLayers imageSubresource;
    VkOffset3D imageOffset; VkExtent3D imageExtent;
} VkBufferImageCopy;
typedef struct VkImageBlit {
    VkImageSubresourceLayers srcSubresource; VkOffset3D srcOffsets[2]; VkImageSubresourceLayers dstSubresource; VkOffset3D dstOffsets[2];
} VkImageBlit;
typedef struct VkClearAttachment { VkFlags aspectMask; uint32_t colorAttachment; VkClearValue clearValue; } VkClearAttachment;
typedef struct VkClearRect { VkRect2D rect; uint32_t baseArrayLayer, layerCount; } VkClearRect;
typedef struct VkDrawIndirectCommand { uint32_t vertexCount, instanceCount, firstVertex, firstInstance; } VkDrawIndirectCommand;
typedef struct VkDrawIndexedIndirectCommand {
    uint32_t indexCount, instanceCount, firstIndex; int32_t vertexOffset; uint32_t firstInstance;
} VkDrawIndexedIndirectCommand;

/* Layout checks: sizes from the specification's member lists (64-bit). */
CPV_VK_STATIC_ASSERT(sizeof(VkApplicationInfo) == 48, "VkApplicationInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkInstanceCreateInfo) == 64, "VkInstanceCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkPhysicalDeviceLimits) == 504, "VkPhysicalDeviceLimits");
CPV_VK_STATIC_ASSERT(sizeof(VkPhysicalDeviceProperties) == 824, "VkPhysicalDeviceProperties");
CPV_VK_STATIC_ASSERT(sizeof(VkPhysicalDeviceFeatures) == 220, "VkPhysicalDeviceFeatures");
CPV_VK_STATIC_ASSERT(sizeof(VkPhysicalDeviceMemoryProperties) == 520, "VkPhysicalDeviceMemoryProperties");
CPV_VK_STATIC_ASSERT(sizeof(VkQueueFamilyProperties) == 24, "VkQueueFamilyProperties");
CPV_VK_STATIC_ASSERT(sizeof(VkDeviceQueueCreateInfo) == 40, "VkDeviceQueueCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkDeviceCreateInfo) == 72, "VkDeviceCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkMemoryAllocateInfo) == 32, "VkMemoryAllocateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkMemoryRequirements) == 24, "VkMemoryRequirements");
CPV_VK_STATIC_ASSERT(sizeof(VkBufferCreateInfo) == 56, "VkBufferCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkImageCreateInfo) == 88, "VkImageCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkSubresourceLayout) == 40, "VkSubresourceLayout");
CPV_VK_STATIC_ASSERT(sizeof(VkImageViewCreateInfo) == 80, "VkImageViewCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkSamplerCreateInfo) == 80, "VkSamplerCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkShaderModuleCreateInfo) == 40, "VkShaderModuleCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkDescriptorSetLayoutBinding) == 24, "VkDescriptorSetLayoutBinding");
CPV_VK_STATIC_ASSERT(sizeof(VkDescriptorSetLayoutCreateInfo) == 32, "VkDescriptorSetLayoutCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkPipelineLayoutCreateInfo) == 48, "VkPipelineLayoutCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkDescriptorPoolCreateInfo) == 40, "VkDescriptorPoolCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkDescriptorSetAllocateInfo) == 40, "VkDescriptorSetAllocateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkDescriptorImageInfo) == 24, "VkDescriptorImageInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkDescriptorBufferInfo) == 24, "VkDescriptorBufferInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkWriteDescriptorSet) == 64, "VkWriteDescriptorSet");
CPV_VK_STATIC_ASSERT(sizeof(VkAttachmentDescription) == 36, "VkAttachmentDescription");
CPV_VK_STATIC_ASSERT(sizeof(VkSubpassDescription) == 72, "VkSubpassDescription");
CPV_VK_STATIC_ASSERT(sizeof(VkRenderPassCreateInfo) == 64, "VkRenderPassCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkFramebufferCreateInfo) == 64, "VkFramebufferCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkPipelineShaderStageCreateInfo) == 48, "VkPipelineShaderStageCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkPipelineVertexInputStateCreateInfo) == 48, "VkPipelineVertexInputStateCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkPipelineInputAssemblyStateCreateInfo) == 32, "VkPipelineInputAssemblyStateCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkPipelineViewportStateCreateInfo) == 48, "VkPipelineViewportStateCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkPipelineRasterizationStateCreateInfo) == 64, "VkPipelineRasterizationStateCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkPipelineMultisampleStateCreateInfo) == 48, "VkPipelineMultisampleStateCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkPipelineDepthStencilStateCreateInfo) == 104, "VkPipelineDepthStencilStateCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkPipelineColorBlendAttachmentState) == 32, "VkPipelineColorBlendAttachmentState");
CPV_VK_STATIC_ASSERT(sizeof(VkPipelineColorBlendStateCreateInfo) == 56, "VkPipelineColorBlendStateCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkPipelineDynamicStateCreateInfo) == 32, "VkPipelineDynamicStateCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkGraphicsPipelineCreateInfo) == 144, "VkGraphicsPipelineCreateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkCommandBufferAllocateInfo) == 32, "VkCommandBufferAllocateInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkCommandBufferBeginInfo) == 32, "VkCommandBufferBeginInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkRenderPassBeginInfo) == 64, "VkRenderPassBeginInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkSubmitInfo) == 72, "VkSubmitInfo");
CPV_VK_STATIC_ASSERT(sizeof(VkFenceCreateInfo) == 24, "VkFenceCreateInfo");
CPV_VK_STATIC_ASSERT(offsetof(VkGraphicsPipelineCreateInfo, layout) == 104, "VkGraphicsPipelineCreateInfo.layout");
CPV_VK_STATIC_ASSERT(sizeof(VkImageCopy) == 68, "VkImageCopy");
CPV_VK_STATIC_ASSERT(sizeof(VkBufferImageCopy) == 56, "VkBufferImageCopy");
CPV_VK_STATIC_ASSERT(sizeof(VkImageBlit) == 80, "VkImageBlit");
CPV_VK_STATIC_ASSERT(sizeof(VkClearAttachment) == 24, "VkClearAttachment");
CPV_VK_STATIC_ASSERT(sizeof(VkClearRect) == 24, "VkClearRect");
CPV_VK_STATIC_ASSERT(sizeof(VkDrawIndexedIndirectCommand) == 20, "VkDrawIndexedIndirectCommand");
CPV_VK_STATIC_ASSERT(offsetof(VkPhysicalDeviceProperties, limits) == 296, "VkPhysicalDeviceProperties.limits");

/* Entry points of the ICD interface (CPVulkan/CPVulkan.cpp:14-104 is the
 * reference's implementation of the same three). */
typedef PFN_vkVoidFunction (VKAPI_PTR* PFN_vkGetInstanceProcAddr)(VkInstance instance, const char* pName);
VkResult vk_icdNegotiateLoaderICDInterfaceVersion(uint32_t* pSupportedVersion);
PFN_vkVoidFunction vk_icdGetInstanceProcAddr(VkInstance instance, const char* pName);
PFN_vkVoidFunction vk_icdGetPhysicalDeviceProcAddr(VkInstance instance, const char* pName);

#ifdef __cplusplus
}
#endif

#endif
