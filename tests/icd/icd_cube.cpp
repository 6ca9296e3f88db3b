// A Vulkan application for the ICD test: renders the 15-draw_cube sample's
// frame through the driver's own entry points (as the loader would reach
// them) and writes the colour and depth images.
//
//   icd_cube <icd.so> <dir> <width> <height> [mode]
//
// draw: vkCmdDraw(36, 1, 0, 0); indirect: the same vertices as two
// VkDrawIndirectCommand records read by vkCmdDrawIndirect; midclear:
// vkCmdDraw, then vkCmdClearAttachments of a colour and a depth rectangle
// (colour (1, 0, 0.5, 1), depth 0.5) over [W/4, W/4 + W/2) x [H/4, H/4 + H/3),
// then vkCmdDraw again, all in one render pass; lines: the vertices as a
// LINE_LIST of width 3; dynamic: constant-colour blending and a depth-bounds
// test with dynamic constants and range, drawn twice: constants (0.25, 0.5,
// 0.75, 1) and bounds [0, 1], then (0.8, 0.1, 0.3, 0.5) and [0, 0.995];
// indexed_indirect: a uint16 index buffer 0..35 and two
// VkDrawIndexedIndirectCommand records read by vkCmdDrawIndexedIndirect;
// indirect_count / indexed_indirect_count: vkCmdDrawIndirectCountKHR /
// vkCmdDrawIndexedIndirectCountAMD with three records, maxDrawCount 3 and a
// count buffer holding 2 (the third record, which would draw the whole cube
// again shifted, must be ignored); stencil: a D24_UNORM_S8_UINT attachment,
// stencil test EQUAL / REPLACE with compare mask, write mask and reference
// all dynamic, drawn twice: masks 0x00 / 0xF0 with reference 0xA5, then
// masks 0xF0 / 0x0F with references 0xAF (front) and 0xA3 (back);
// multipass: three framebuffers (own colour and depth images, clear colours
// 0.2, 0.3, 0.4) drawn by three render passes of one command buffer, one
// submit; resubmit: the same command buffer recorded without
// ONE_TIME_SUBMIT and submitted three times, the uniform rewritten from
// mvp0.bin, mvp1.bin, mvp2.bin before each submit, with the fence checks
// printed as "fence <what>: <VkResult>".
//
// <dir> holds vs.spv, fs.spv (SPIR-V), vertices.bin (36 x {vec4 pos, vec4
// colour}) and mvp.bin (mat4); the program writes color.bin (B8G8R8A8) and
// depth.bin (D16, or D24S8 in stencil mode) there; multipass and resubmit
// write color_<submit>_<pass>.bin and depth_<submit>_<pass>.bin.  Pipeline state as the sample: LESS_OR_EQUAL depth,
// FRONT_FACE_CLOCKWISE, cull NONE, clear colour 0.2, depth 1.0, dynamic
// viewport (Samples/15-draw_cube/15-draw_cube.cpp:38-149,
// Samples/utils/util_init.cpp:1541-1693).
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

static std::vector<char> read_file(const std::string& p)
{
    std::ifstream f(p, std::ios::binary);
    if (!f) {
        std::fprintf(stderr, "cannot read %s\n", p.c_str());
        std::exit(2);
    }
    return std::vector<char>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

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
    if (argc != 5 && argc != 6) {
        std::fprintf(stderr, "usage: icd_cube <icd.so> <dir> <width> <height> [mode]\n");
        return 2;
    }
    const std::string mode = argc == 6 ? argv[5] : "draw";
    const std::string dir = argv[2];
    const uint32_t W = (uint32_t)std::atoi(argv[3]), H = (uint32_t)std::atoi(argv[4]);
    void* lib = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if (!lib) {
        std::fprintf(stderr, "dlopen: %s\n", dlerror());
        return 2;
    }
    auto negotiate = reinterpret_cast<VkResult (*)(uint32_t*)>(dlsym(lib, "vk_icdNegotiateLoaderICDInterfaceVersion"));
    gipa = reinterpret_cast<PFN_vkGetInstanceProcAddr>(dlsym(lib, "vk_icdGetInstanceProcAddr"));
    if (!negotiate || !gipa) return 2;
    uint32_t ver = 6;
    CHECK(negotiate(&ver));
    if (ver != 5) return 3;

    auto vkCreateInstance = reinterpret_cast<VkResult (*)(const VkInstanceCreateInfo*, const VkAllocationCallbacks*, VkInstance*)>(
        gipa(nullptr, "vkCreateInstance"));
    VkApplicationInfo app{VK_STRUCTURE_TYPE_APPLICATION_INFO, nullptr, "icd_cube", 1, "none", 1, VK_API_VERSION_1_1};
    VkInstanceCreateInfo ici{VK_STRUCTURE_TYPE_INSTANCE_CREATE_INFO, nullptr, 0, &app, 0, nullptr, 0, nullptr};
    CHECK(vkCreateInstance(&ici, nullptr, &g_inst));

    FN(EnumeratePhysicalDevices, VkInstance, uint32_t*, VkPhysicalDevice*);
    VFN(GetPhysicalDeviceProperties, VkPhysicalDevice, VkPhysicalDeviceProperties*);
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
    FN(CreateImageView, VkDevice, const VkImageViewCreateInfo*, const VkAllocationCallbacks*, VkImageView*);
    FN(CreateShaderModule, VkDevice, const VkShaderModuleCreateInfo*, const VkAllocationCallbacks*, VkShaderModule*);
    FN(CreateDescriptorSetLayout, VkDevice, const VkDescriptorSetLayoutCreateInfo*, const VkAllocationCallbacks*, VkDescriptorSetLayout*);
    FN(CreatePipelineLayout, VkDevice, const VkPipelineLayoutCreateInfo*, const VkAllocationCallbacks*, VkPipelineLayout*);
    FN(CreateDescriptorPool, VkDevice, const VkDescriptorPoolCreateInfo*, const VkAllocationCallbacks*, VkDescriptorPool*);
    FN(AllocateDescriptorSets, VkDevice, const VkDescriptorSetAllocateInfo*, VkDescriptorSet*);
    VFN(UpdateDescriptorSets, VkDevice, uint32_t, const VkWriteDescriptorSet*, uint32_t, const VkCopyDescriptorSet*);
    FN(CreateRenderPass, VkDevice, const VkRenderPassCreateInfo*, const VkAllocationCallbacks*, VkRenderPass*);
    FN(CreateFramebuffer, VkDevice, const VkFramebufferCreateInfo*, const VkAllocationCallbacks*, VkFramebuffer*);
    FN(CreateGraphicsPipelines, VkDevice, VkPipelineCache, uint32_t, const VkGraphicsPipelineCreateInfo*, const VkAllocationCallbacks*,
       VkPipeline*);
    FN(CreateCommandPool, VkDevice, const VkCommandPoolCreateInfo*, const VkAllocationCallbacks*, VkCommandPool*);
    FN(AllocateCommandBuffers, VkDevice, const VkCommandBufferAllocateInfo*, VkCommandBuffer*);
    FN(BeginCommandBuffer, VkCommandBuffer, const VkCommandBufferBeginInfo*);
    FN(EndCommandBuffer, VkCommandBuffer);
    VFN(CmdBeginRenderPass, VkCommandBuffer, const VkRenderPassBeginInfo*, int32_t);
    VFN(CmdEndRenderPass, VkCommandBuffer);
    VFN(CmdBindPipeline, VkCommandBuffer, int32_t, VkPipeline);
    VFN(CmdBindVertexBuffers, VkCommandBuffer, uint32_t, uint32_t, const VkBuffer*, const VkDeviceSize*);
    VFN(CmdBindDescriptorSets, VkCommandBuffer, int32_t, VkPipelineLayout, uint32_t, uint32_t, const VkDescriptorSet*, uint32_t,
        const uint32_t*);
    VFN(CmdSetViewport, VkCommandBuffer, uint32_t, uint32_t, const VkViewport*);
    VFN(CmdSetScissor, VkCommandBuffer, uint32_t, uint32_t, const VkRect2D*);
    VFN(CmdSetBlendConstants, VkCommandBuffer, const float*);
    VFN(CmdSetDepthBounds, VkCommandBuffer, float, float);
    VFN(CmdDraw, VkCommandBuffer, uint32_t, uint32_t, uint32_t, uint32_t);
    VFN(CmdDrawIndirect, VkCommandBuffer, VkBuffer, VkDeviceSize, uint32_t, uint32_t);
    VFN(CmdDrawIndexedIndirect, VkCommandBuffer, VkBuffer, VkDeviceSize, uint32_t, uint32_t);
    VFN(CmdBindIndexBuffer, VkCommandBuffer, VkBuffer, VkDeviceSize, int32_t);
    VFN(CmdSetStencilCompareMask, VkCommandBuffer, VkFlags, uint32_t);
    VFN(CmdSetStencilWriteMask, VkCommandBuffer, VkFlags, uint32_t);
    VFN(CmdSetStencilReference, VkCommandBuffer, VkFlags, uint32_t);
    FN(GetFenceStatus, VkDevice, VkFence);
    FN(ResetFences, VkDevice, uint32_t, const VkFence*);
    FN(EnumerateDeviceExtensionProperties, VkPhysicalDevice, const char*, uint32_t*, VkExtensionProperties*);
    VFN(CmdClearAttachments, VkCommandBuffer, uint32_t, const VkClearAttachment*, uint32_t, const VkClearRect*);
    FN(QueueSubmit, VkQueue, uint32_t, const VkSubmitInfo*, VkFence);
    FN(CreateFence, VkDevice, const VkFenceCreateInfo*, const VkAllocationCallbacks*, VkFence*);
    FN(WaitForFences, VkDevice, uint32_t, const VkFence*, VkBool32, uint64_t);
    VFN(DestroyDevice, VkDevice, const VkAllocationCallbacks*);
    VFN(DestroyInstance, VkInstance, const VkAllocationCallbacks*);

    uint32_t n = 1;
    VkPhysicalDevice phys;
    CHECK(vkEnumeratePhysicalDevices(g_inst, &n, &phys));
    VkPhysicalDeviceProperties props;
    vkGetPhysicalDeviceProperties(phys, &props);
    std::printf("device: %s\n", props.deviceName);
    VkPhysicalDeviceMemoryProperties mp;
    vkGetPhysicalDeviceMemoryProperties(phys, &mp);
    uint32_t host_type = ~0u;
    for (uint32_t i = 0; i < mp.memoryTypeCount; i++)
        if (mp.memoryTypes[i].propertyFlags & VK_MEMORY_PROPERTY_HOST_VISIBLE_BIT) { host_type = i; break; }
    if (host_type == ~0u) return 4;

    // the indirect-count draws are device extensions
    uint32_t n_ext = 0;
    CHECK(vkEnumerateDeviceExtensionProperties(phys, nullptr, &n_ext, nullptr));
    std::vector<VkExtensionProperties> exts(n_ext);
    CHECK(vkEnumerateDeviceExtensionProperties(phys, nullptr, &n_ext, exts.data()));
    std::vector<const char*> ext_names;
    for (const VkExtensionProperties& e : exts) ext_names.push_back(e.extensionName);
    const float prio = 1.0f;
    VkDeviceQueueCreateInfo qci{VK_STRUCTURE_TYPE_DEVICE_QUEUE_CREATE_INFO, nullptr, 0, 0, 1, &prio};
    VkDeviceCreateInfo dci{VK_STRUCTURE_TYPE_DEVICE_CREATE_INFO, nullptr, 0, 1, &qci, 0, nullptr, (uint32_t)ext_names.size(),
                           ext_names.data(), nullptr};
    VkDevice dev;
    CHECK(vkCreateDevice(phys, &dci, nullptr, &dev));
    VkQueue queue;
    vkGetDeviceQueue(dev, 0, 0, &queue);

    auto host_memory = [&](VkMemoryRequirements req, VkDeviceMemory* mem) {
        VkMemoryAllocateInfo ai{VK_STRUCTURE_TYPE_MEMORY_ALLOCATE_INFO, nullptr, req.size, host_type};
        CHECK(vkAllocateMemory(dev, &ai, nullptr, mem));
    };
    auto buffer = [&](const std::vector<char>& data, VkBuffer* buf) -> void* {
        VkBufferCreateInfo bi{VK_STRUCTURE_TYPE_BUFFER_CREATE_INFO, nullptr, 0, data.size(), 0x1D0, 0, 0, nullptr};
        CHECK(vkCreateBuffer(dev, &bi, nullptr, buf));
        VkMemoryRequirements req;
        vkGetBufferMemoryRequirements(dev, *buf, &req);
        VkDeviceMemory mem;
        host_memory(req, &mem);
        void* p;
        CHECK(vkMapMemory(dev, mem, 0, req.size, 0, &p));
        std::memcpy(p, data.data(), data.size());
        CHECK(vkBindBufferMemory(dev, *buf, mem, 0));
        return p;
    };
    auto image = [&](int32_t format, VkImage* img, VkImageView* view, void** mapped) {
        VkImageCreateInfo ii{VK_STRUCTURE_TYPE_IMAGE_CREATE_INFO, nullptr, 0, VK_IMAGE_TYPE_2D, format, {W, H, 1}, 1, 1, 1,
                             VK_IMAGE_TILING_LINEAR, 0x10 | 0x20, 0, 0, nullptr, 0};
        CHECK(vkCreateImage(dev, &ii, nullptr, img));
        VkMemoryRequirements req;
        vkGetImageMemoryRequirements(dev, *img, &req);
        VkDeviceMemory mem;
        host_memory(req, &mem);
        CHECK(vkMapMemory(dev, mem, 0, req.size, 0, mapped));
        CHECK(vkBindImageMemory(dev, *img, mem, 0));
        VkImageViewCreateInfo vi{VK_STRUCTURE_TYPE_IMAGE_VIEW_CREATE_INFO, nullptr, 0, *img, 1, format, {0, 0, 0, 0}, {1, 0, 1, 0, 1}};
        CHECK(vkCreateImageView(dev, &vi, nullptr, view));
    };

    const bool stencil = mode == "stencil";
    const bool many = mode == "multipass" || mode == "resubmit";
    const uint32_t passes = many ? 3 : 1, submits = mode == "resubmit" ? 3 : 1;
    const int32_t ds_format = stencil ? CPV_FORMAT_D24_UNORM_S8_UINT : CPV_FORMAT_D16_UNORM;
    const uint32_t ds_bytes = stencil ? 4 : 2;
    VkBuffer vbuf, ubuf;
    buffer(read_file(dir + "/vertices.bin"), &vbuf);
    void* mvp_map = buffer(read_file(dir + (mode == "resubmit" ? "/mvp0.bin" : "/mvp.bin")), &ubuf);
    std::vector<VkImage> color(passes), depth(passes);
    std::vector<VkImageView> color_view(passes), depth_view(passes);
    std::vector<void*> color_map(passes), depth_map(passes);
    for (uint32_t k = 0; k < passes; k++) {
        image(CPV_FORMAT_B8G8R8A8_UNORM, &color[k], &color_view[k], &color_map[k]);
        image(ds_format, &depth[k], &depth_view[k], &depth_map[k]);
    }

    VkDescriptorSetLayoutBinding lb{0, VK_DESCRIPTOR_TYPE_UNIFORM_BUFFER, 1, VK_SHADER_STAGE_VERTEX_BIT, nullptr};
    VkDescriptorSetLayoutCreateInfo dli{VK_STRUCTURE_TYPE_DESCRIPTOR_SET_LAYOUT_CREATE_INFO, nullptr, 0, 1, &lb};
    VkDescriptorSetLayout dsl;
    CHECK(vkCreateDescriptorSetLayout(dev, &dli, nullptr, &dsl));
    VkPipelineLayoutCreateInfo pli{VK_STRUCTURE_TYPE_PIPELINE_LAYOUT_CREATE_INFO, nullptr, 0, 1, &dsl, 0, nullptr};
    VkPipelineLayout layout;
    CHECK(vkCreatePipelineLayout(dev, &pli, nullptr, &layout));
    VkDescriptorPoolSize ps{VK_DESCRIPTOR_TYPE_UNIFORM_BUFFER, 1};
    VkDescriptorPoolCreateInfo dpi{VK_STRUCTURE_TYPE_DESCRIPTOR_POOL_CREATE_INFO, nullptr, 0, 1, 1, &ps};
    VkDescriptorPool pool;
    CHECK(vkCreateDescriptorPool(dev, &dpi, nullptr, &pool));
    VkDescriptorSetAllocateInfo dai{VK_STRUCTURE_TYPE_DESCRIPTOR_SET_ALLOCATE_INFO, nullptr, pool, 1, &dsl};
    VkDescriptorSet set;
    CHECK(vkAllocateDescriptorSets(dev, &dai, &set));
    VkDescriptorBufferInfo dbi{ubuf, 0, 64};
    VkWriteDescriptorSet wr{VK_STRUCTURE_TYPE_WRITE_DESCRIPTOR_SET, nullptr, set, 0, 0, 1, VK_DESCRIPTOR_TYPE_UNIFORM_BUFFER, nullptr,
                            &dbi, nullptr};
    vkUpdateDescriptorSets(dev, 1, &wr, 0, nullptr);

    VkAttachmentDescription att[2] = {
        {0, CPV_FORMAT_B8G8R8A8_UNORM, 1, VK_ATTACHMENT_LOAD_OP_CLEAR, 0, 2, 1, 0, 2},
        {0, ds_format, 1, VK_ATTACHMENT_LOAD_OP_CLEAR, 1, VK_ATTACHMENT_LOAD_OP_CLEAR, 1, 0, 3},
    };
    VkAttachmentReference cref{0, 2}, dref{1, 3};
    VkSubpassDescription sp{0, VK_PIPELINE_BIND_POINT_GRAPHICS, 0, nullptr, 1, &cref, nullptr, &dref, 0, nullptr};
    VkRenderPassCreateInfo rpi{VK_STRUCTURE_TYPE_RENDER_PASS_CREATE_INFO, nullptr, 0, 2, att, 1, &sp, 0, nullptr};
    VkRenderPass rp;
    CHECK(vkCreateRenderPass(dev, &rpi, nullptr, &rp));
    std::vector<VkFramebuffer> fbs(passes);
    for (uint32_t k = 0; k < passes; k++) {
        VkImageView views[2] = {color_view[k], depth_view[k]};
        VkFramebufferCreateInfo fbi{VK_STRUCTURE_TYPE_FRAMEBUFFER_CREATE_INFO, nullptr, 0, rp, 2, views, W, H, 1};
        CHECK(vkCreateFramebuffer(dev, &fbi, nullptr, &fbs[k]));
    }

    std::vector<char> vs_code = read_file(dir + "/vs.spv"), fs_code = read_file(dir + "/fs.spv");
    VkShaderModuleCreateInfo smi{VK_STRUCTURE_TYPE_SHADER_MODULE_CREATE_INFO, nullptr, 0, vs_code.size(),
                                 reinterpret_cast<const uint32_t*>(vs_code.data())};
    VkShaderModule vsm, fsm;
    CHECK(vkCreateShaderModule(dev, &smi, nullptr, &vsm));
    smi.codeSize = fs_code.size();
    smi.pCode = reinterpret_cast<const uint32_t*>(fs_code.data());
    CHECK(vkCreateShaderModule(dev, &smi, nullptr, &fsm));
    VkPipelineShaderStageCreateInfo stages[2] = {
        {VK_STRUCTURE_TYPE_PIPELINE_SHADER_STAGE_CREATE_INFO, nullptr, 0, VK_SHADER_STAGE_VERTEX_BIT, vsm, "main", nullptr},
        {VK_STRUCTURE_TYPE_PIPELINE_SHADER_STAGE_CREATE_INFO, nullptr, 0, VK_SHADER_STAGE_FRAGMENT_BIT, fsm, "main", nullptr},
    };
    VkVertexInputBindingDescription vb{0, 32, 0};
    VkVertexInputAttributeDescription va[2] = {{0, 0, CPV_FORMAT_R32G32B32A32_SFLOAT, 0}, {1, 0, CPV_FORMAT_R32G32B32A32_SFLOAT, 16}};
    VkPipelineVertexInputStateCreateInfo vis{VK_STRUCTURE_TYPE_PIPELINE_VERTEX_INPUT_STATE_CREATE_INFO, nullptr, 0, 1, &vb, 2, va};
    VkPipelineInputAssemblyStateCreateInfo ias{VK_STRUCTURE_TYPE_PIPELINE_INPUT_ASSEMBLY_STATE_CREATE_INFO, nullptr, 0,
                                               CPV_TOPOLOGY_TRIANGLE_LIST, 0};
    VkPipelineViewportStateCreateInfo vps{VK_STRUCTURE_TYPE_PIPELINE_VIEWPORT_STATE_CREATE_INFO, nullptr, 0, 1, nullptr, 1, nullptr};
    VkPipelineRasterizationStateCreateInfo rs{VK_STRUCTURE_TYPE_PIPELINE_RASTERIZATION_STATE_CREATE_INFO, nullptr, 0, 0, 0,
                                              CPV_POLYGON_FILL, CPV_CULL_NONE, CPV_FRONT_FACE_CW, 0, 0, 0, 0, 1.0f};
    VkPipelineMultisampleStateCreateInfo ms{VK_STRUCTURE_TYPE_PIPELINE_MULTISAMPLE_STATE_CREATE_INFO, nullptr, 0, 1, 0, 0, nullptr, 0, 0};
    VkStencilOpState keep{0, 0, 0, CPV_COMPARE_ALWAYS, 0, 0, 0};
    VkPipelineDepthStencilStateCreateInfo dss{VK_STRUCTURE_TYPE_PIPELINE_DEPTH_STENCIL_STATE_CREATE_INFO, nullptr, 0, 1, 1,
                                              CPV_COMPARE_LESS_OR_EQUAL, 0, 0, keep, keep, 0.0f, 0.0f};
    VkPipelineColorBlendAttachmentState cba{0, 0, 0, 0, 0, 0, 0, 0xF};
    VkPipelineColorBlendStateCreateInfo cbs{VK_STRUCTURE_TYPE_PIPELINE_COLOR_BLEND_STATE_CREATE_INFO, nullptr, 0, 0, 0, 1, &cba,
                                            {1.0f, 1.0f, 1.0f, 1.0f}};
    if (mode == "lines") {
        // the cube's 36 vertices as 18 lines, 3 pixels wide
        ias.topology = CPV_TOPOLOGY_LINE_LIST;
        rs.lineWidth = 3.0f;
    }
    int32_t dyn[5] = {VK_DYNAMIC_STATE_VIEWPORT, VK_DYNAMIC_STATE_SCISSOR, VK_DYNAMIC_STATE_BLEND_CONSTANTS,
                      VK_DYNAMIC_STATE_DEPTH_BOUNDS, 0};
    VkPipelineDynamicStateCreateInfo dys{VK_STRUCTURE_TYPE_PIPELINE_DYNAMIC_STATE_CREATE_INFO, nullptr, 0, 2, dyn};
    int32_t dyn_stencil[5] = {VK_DYNAMIC_STATE_VIEWPORT, VK_DYNAMIC_STATE_SCISSOR, VK_DYNAMIC_STATE_STENCIL_COMPARE_MASK,
                              VK_DYNAMIC_STATE_STENCIL_WRITE_MASK, VK_DYNAMIC_STATE_STENCIL_REFERENCE};
    if (stencil) {
        // EQUAL / REPLACE on both faces; the pipeline's masks and reference
        // (all zero) are replaced by vkCmdSetStencil*
        const VkStencilOpState st{CPV_STENCIL_KEEP, CPV_STENCIL_REPLACE, CPV_STENCIL_KEEP, CPV_COMPARE_EQUAL, 0, 0, 0};
        dss.stencilTestEnable = 1;
        dss.front = dss.back = st;
        dys.pDynamicStates = dyn_stencil;
        dys.dynamicStateCount = 5;
    }
    if (mode == "dynamic") {
        // constant-colour blending and a depth-bounds test whose constants
        // and range come from vkCmdSetBlendConstants / vkCmdSetDepthBounds
        // (the pipeline's own values would reject every fragment)
        cba = VkPipelineColorBlendAttachmentState{1, 10, 11, 0, 12, 13, 0, 0xF};   // CONSTANT_COLOR, ONE_MINUS_..., CONSTANT_ALPHA, ...
        dss.depthBoundsTestEnable = 1;
        dss.minDepthBounds = dss.maxDepthBounds = -1.0f;
        dys.dynamicStateCount = 4;
    }
    VkGraphicsPipelineCreateInfo gpi{VK_STRUCTURE_TYPE_GRAPHICS_PIPELINE_CREATE_INFO, nullptr, 0, 2, stages, &vis, &ias, nullptr, &vps,
                                     &rs, &ms, &dss, &cbs, &dys, layout, rp, 0, nullptr, -1};
    VkPipeline pipe;
    CHECK(vkCreateGraphicsPipelines(dev, nullptr, 1, &gpi, nullptr, &pipe));

    VkCommandPoolCreateInfo cpi{VK_STRUCTURE_TYPE_COMMAND_POOL_CREATE_INFO, nullptr, 0, 0};
    VkCommandPool cpool;
    CHECK(vkCreateCommandPool(dev, &cpi, nullptr, &cpool));
    VkCommandBufferAllocateInfo cai{VK_STRUCTURE_TYPE_COMMAND_BUFFER_ALLOCATE_INFO, nullptr, cpool, VK_COMMAND_BUFFER_LEVEL_PRIMARY, 1};
    VkCommandBuffer cmd;
    CHECK(vkAllocateCommandBuffers(dev, &cai, &cmd));
    // a command buffer submitted more than once is recorded without ONE_TIME_SUBMIT
    VkCommandBufferBeginInfo cbi{VK_STRUCTURE_TYPE_COMMAND_BUFFER_BEGIN_INFO, nullptr,
                                 mode == "resubmit" ? 0u : (VkFlags)VK_COMMAND_BUFFER_USAGE_ONE_TIME_SUBMIT_BIT, nullptr};
    CHECK(vkBeginCommandBuffer(cmd, &cbi));
    for (uint32_t pass = 0; pass < passes; pass++) {
    VkClearValue clears[2];
    for (int i = 0; i < 4; i++) clears[0].color.float32[i] = 0.2f + 0.1f * (float)pass;
    clears[1].depthStencil = VkClearDepthStencilValue{1.0f, 0};
    VkRenderPassBeginInfo rbi{VK_STRUCTURE_TYPE_RENDER_PASS_BEGIN_INFO, nullptr, rp, fbs[pass], {{0, 0}, {W, H}}, 2, clears};
    vkCmdBeginRenderPass(cmd, &rbi, VK_SUBPASS_CONTENTS_INLINE);
    vkCmdBindPipeline(cmd, VK_PIPELINE_BIND_POINT_GRAPHICS, pipe);
    vkCmdBindDescriptorSets(cmd, VK_PIPELINE_BIND_POINT_GRAPHICS, layout, 0, 1, &set, 0, nullptr);
    VkViewport vp{0.0f, 0.0f, (float)W, (float)H, 0.0f, 1.0f};
    vkCmdSetViewport(cmd, 0, 1, &vp);
    VkRect2D sc{{0, 0}, {W, H}};
    vkCmdSetScissor(cmd, 0, 1, &sc);
    const VkDeviceSize off = 0;
    vkCmdBindVertexBuffers(cmd, 0, 1, &vbuf, &off);
    VkBuffer ibuf = nullptr;
    auto bytes_of = [](const void* p, size_t n) { return std::vector<char>((const char*)p, (const char*)p + n); };
    if (mode == "indirect") {
        const VkDrawIndirectCommand recs[2] = {{18, 1, 0, 0}, {18, 1, 18, 0}};
        buffer(bytes_of(recs, sizeof(recs)), &ibuf);
        vkCmdDrawIndirect(cmd, ibuf, 0, 2, sizeof(VkDrawIndirectCommand));
    } else if (mode == "indexed_indirect" || mode == "indexed_indirect_count") {
        uint16_t idx[36];
        for (int i = 0; i < 36; i++) idx[i] = (uint16_t)i;
        VkBuffer index_buf;
        buffer(bytes_of(idx, sizeof(idx)), &index_buf);
        vkCmdBindIndexBuffer(cmd, index_buf, 0, VK_INDEX_TYPE_UINT16);
        // the third record (all 36 indices, vertexOffset 1) is past the count
        const VkDrawIndexedIndirectCommand recs[3] = {{18, 1, 0, 0, 0}, {18, 1, 18, 0, 0}, {36, 1, 0, 1, 0}};
        buffer(bytes_of(recs, sizeof(recs)), &ibuf);
        if (mode == "indexed_indirect") {
            vkCmdDrawIndexedIndirect(cmd, ibuf, 0, 2, sizeof(VkDrawIndexedIndirectCommand));
        } else {
            const uint32_t count = 2;
            VkBuffer cbuf;
            buffer(bytes_of(&count, 4), &cbuf);
            auto draw = load<void (*)(VkCommandBuffer, VkBuffer, VkDeviceSize, VkBuffer, VkDeviceSize, uint32_t, uint32_t)>(
                "vkCmdDrawIndexedIndirectCountAMD");
            draw(cmd, ibuf, 0, cbuf, 0, 3, sizeof(VkDrawIndexedIndirectCommand));
        }
    } else if (mode == "indirect_count") {
        // records 24 bytes apart, with a third one past the count
        struct Rec {
            VkDrawIndirectCommand c;
            uint32_t pad[2];
        };
        const Rec recs[3] = {{{18, 1, 0, 0}, {7, 7}}, {{18, 1, 18, 0}, {7, 7}}, {{36, 2, 0, 0}, {7, 7}}};
        const uint32_t count[2] = {0xFFFFFFFFu, 2};   // read at offset 4
        buffer(bytes_of(recs, sizeof(recs)), &ibuf);
        VkBuffer cbuf;
        buffer(bytes_of(count, sizeof(count)), &cbuf);
        auto draw = load<void (*)(VkCommandBuffer, VkBuffer, VkDeviceSize, VkBuffer, VkDeviceSize, uint32_t, uint32_t)>(
            "vkCmdDrawIndirectCountKHR");
        draw(cmd, ibuf, 0, cbuf, 4, 3, sizeof(Rec));
    } else if (stencil) {
        vkCmdSetStencilCompareMask(cmd, VK_STENCIL_FACE_FRONT_AND_BACK, 0x00);
        vkCmdSetStencilWriteMask(cmd, VK_STENCIL_FACE_FRONT_AND_BACK, 0xF0);
        vkCmdSetStencilReference(cmd, VK_STENCIL_FACE_FRONT_AND_BACK, 0xA5);
        vkCmdDraw(cmd, 36, 1, 0, 0);
        vkCmdSetStencilCompareMask(cmd, VK_STENCIL_FACE_FRONT_AND_BACK, 0xF0);
        vkCmdSetStencilWriteMask(cmd, VK_STENCIL_FACE_FRONT_AND_BACK, 0x0F);
        vkCmdSetStencilReference(cmd, VK_STENCIL_FACE_FRONT_BIT, 0xAF);
        vkCmdSetStencilReference(cmd, VK_STENCIL_FACE_BACK_BIT, 0xA3);
        vkCmdDraw(cmd, 36, 1, 0, 0);
    } else if (mode == "dynamic") {
        const float k1[4] = {0.25f, 0.5f, 0.75f, 1.0f}, k2[4] = {0.8f, 0.1f, 0.3f, 0.5f};
        vkCmdSetBlendConstants(cmd, k1);
        vkCmdSetDepthBounds(cmd, 0.0f, 1.0f);
        vkCmdDraw(cmd, 36, 1, 0, 0);
        vkCmdSetBlendConstants(cmd, k2);
        vkCmdSetDepthBounds(cmd, 0.0f, 0.995f);
        vkCmdDraw(cmd, 36, 1, 0, 0);
    } else {
        vkCmdDraw(cmd, 36, 1, 0, 0);   // 15-draw_cube.cpp:149
    }
    if (mode == "midclear") {
        VkClearAttachment ca[2];
        ca[0].aspectMask = VK_IMAGE_ASPECT_COLOR_BIT;
        ca[0].colorAttachment = 0;
        ca[0].clearValue.color.float32[0] = 1.0f;
        ca[0].clearValue.color.float32[1] = 0.0f;
        ca[0].clearValue.color.float32[2] = 0.5f;
        ca[0].clearValue.color.float32[3] = 1.0f;
        ca[1].aspectMask = VK_IMAGE_ASPECT_DEPTH_BIT;
        ca[1].colorAttachment = 0;
        ca[1].clearValue.depthStencil = VkClearDepthStencilValue{0.5f, 0};
        VkClearRect cr{{{(int32_t)(W / 4), (int32_t)(H / 4)}, {W / 2, H / 3}}, 0, 1};
        vkCmdClearAttachments(cmd, 2, ca, 1, &cr);
        vkCmdDraw(cmd, 36, 1, 0, 0);
    }
    vkCmdEndRenderPass(cmd);
    }
    CHECK(vkEndCommandBuffer(cmd));

    VkFenceCreateInfo fci{VK_STRUCTURE_TYPE_FENCE_CREATE_INFO, nullptr, 0};
    VkFence fence;
    CHECK(vkCreateFence(dev, &fci, nullptr, &fence));
    VkSubmitInfo si{VK_STRUCTURE_TYPE_SUBMIT_INFO, nullptr, 0, nullptr, nullptr, 1, &cmd, 0, nullptr};
    if (mode == "resubmit") {
        // a fence never submitted: not ready, and a timed wait times out;
        // with a signalled fence beside it, waiting for any one succeeds
        VkFenceCreateInfo on{VK_STRUCTURE_TYPE_FENCE_CREATE_INFO, nullptr, VK_FENCE_CREATE_SIGNALED_BIT};
        VkFence lit;
        CHECK(vkCreateFence(dev, &on, nullptr, &lit));
        const VkFence both[2] = {fence, lit};
        std::printf("fence unsubmitted status: %d\n", vkGetFenceStatus(dev, fence));
        std::printf("fence unsubmitted wait 1ms: %d\n", vkWaitForFences(dev, 1, &fence, VK_TRUE, 1000000ull));
        std::printf("fence any of two: %d\n", vkWaitForFences(dev, 2, both, VK_FALSE, 0));
        std::printf("fence all of two: %d\n", vkWaitForFences(dev, 2, both, VK_TRUE, 0));
    }
    for (uint32_t sub = 0; sub < submits; sub++) {
        if (sub > 0) {
            // the previous submit has finished (its fence waited on): rewrite the uniform
            std::vector<char> m = read_file(dir + "/mvp" + std::to_string(sub) + ".bin");
            std::memcpy(mvp_map, m.data(), m.size());
            CHECK(vkResetFences(dev, 1, &fence));
            std::printf("fence after reset: %d\n", vkGetFenceStatus(dev, fence));
        }
        CHECK(vkQueueSubmit(queue, 1, &si, fence));
        CHECK(vkWaitForFences(dev, 1, &fence, VK_TRUE, ~0ull));
        if (mode == "resubmit") std::printf("fence after wait: %d\n", vkGetFenceStatus(dev, fence));
        for (uint32_t k = 0; k < passes; k++) {
            const std::string tag = many ? "_" + std::to_string(sub) + "_" + std::to_string(k) : "";
            std::ofstream(dir + "/color" + tag + ".bin", std::ios::binary).write((const char*)color_map[k], (std::streamsize)W * H * 4);
            std::ofstream(dir + "/depth" + tag + ".bin", std::ios::binary)
                .write((const char*)depth_map[k], (std::streamsize)W * H * ds_bytes);
        }
    }
    vkDestroyDevice(dev, nullptr);
    vkDestroyInstance(g_inst, nullptr);
    std::printf("rendered %ux%u\n", W, H);
    return 0;
}
