// A Vulkan application for the ICD test: renders the 15-draw_cube sample's
// frame through the driver's own entry points (as the loader would reach
// them) and writes the colour and depth images.
//
//   icd_cube <icd.so> <dir> <width> <height> [draw|indirect|midclear|lines|dynamic]
//
// draw: vkCmdDraw(36, 1, 0, 0); indirect: the same vertices as two
// VkDrawIndirectCommand records read by vkCmdDrawIndirect; midclear:
// vkCmdDraw, then vkCmdClearAttachments of a colour and a depth rectangle
// (colour (1, 0, 0.5, 1), depth 0.5) over [W/4, W/4 + W/2) x [H/4, H/4 + H/3),
// then vkCmdDraw again, all in one render pass; lines: the vertices as a
// LINE_LIST of width 3; dynamic: constant-colour blending and a depth-bounds
// test with dynamic constants and range, drawn twice: constants (0.25, 0.5,
// 0.75, 1) and bounds [0, 1], then (0.8, 0.1, 0.3, 0.5) and [0, 0.995].
//
// <dir> holds vs.spv, fs.spv (SPIR-V), vertices.bin (36 x {vec4 pos, vec4
// colour}) and mvp.bin (mat4); the program writes color.bin (B8G8R8A8) and
// depth.bin (D16) there.  Pipeline state as the sample: LESS_OR_EQUAL depth,
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
        std::fprintf(stderr, "usage: icd_cube <icd.so> <dir> <width> <height> [draw|indirect|midclear]\n");
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

    const float prio = 1.0f;
    VkDeviceQueueCreateInfo qci{VK_STRUCTURE_TYPE_DEVICE_QUEUE_CREATE_INFO, nullptr, 0, 0, 1, &prio};
    VkDeviceCreateInfo dci{VK_STRUCTURE_TYPE_DEVICE_CREATE_INFO, nullptr, 0, 1, &qci, 0, nullptr, 0, nullptr, nullptr};
    VkDevice dev;
    CHECK(vkCreateDevice(phys, &dci, nullptr, &dev));
    VkQueue queue;
    vkGetDeviceQueue(dev, 0, 0, &queue);

    auto host_memory = [&](VkMemoryRequirements req, VkDeviceMemory* mem) {
        VkMemoryAllocateInfo ai{VK_STRUCTURE_TYPE_MEMORY_ALLOCATE_INFO, nullptr, req.size, host_type};
        CHECK(vkAllocateMemory(dev, &ai, nullptr, mem));
    };
    auto buffer = [&](const std::vector<char>& data, VkBuffer* buf) {
        VkBufferCreateInfo bi{VK_STRUCTURE_TYPE_BUFFER_CREATE_INFO, nullptr, 0, data.size(), 0x90, 0, 0, nullptr};
        CHECK(vkCreateBuffer(dev, &bi, nullptr, buf));
        VkMemoryRequirements req;
        vkGetBufferMemoryRequirements(dev, *buf, &req);
        VkDeviceMemory mem;
        host_memory(req, &mem);
        void* p;
        CHECK(vkMapMemory(dev, mem, 0, req.size, 0, &p));
        std::memcpy(p, data.data(), data.size());
        CHECK(vkBindBufferMemory(dev, *buf, mem, 0));
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

    VkBuffer vbuf, ubuf;
    buffer(read_file(dir + "/vertices.bin"), &vbuf);
    buffer(read_file(dir + "/mvp.bin"), &ubuf);
    VkImage color, depth;
    VkImageView color_view, depth_view;
    void *color_map, *depth_map;
    image(CPV_FORMAT_B8G8R8A8_UNORM, &color, &color_view, &color_map);
    image(CPV_FORMAT_D16_UNORM, &depth, &depth_view, &depth_map);

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
        {0, CPV_FORMAT_D16_UNORM, 1, VK_ATTACHMENT_LOAD_OP_CLEAR, 1, 2, 1, 0, 3},
    };
    VkAttachmentReference cref{0, 2}, dref{1, 3};
    VkSubpassDescription sp{0, VK_PIPELINE_BIND_POINT_GRAPHICS, 0, nullptr, 1, &cref, nullptr, &dref, 0, nullptr};
    VkRenderPassCreateInfo rpi{VK_STRUCTURE_TYPE_RENDER_PASS_CREATE_INFO, nullptr, 0, 2, att, 1, &sp, 0, nullptr};
    VkRenderPass rp;
    CHECK(vkCreateRenderPass(dev, &rpi, nullptr, &rp));
    VkImageView views[2] = {color_view, depth_view};
    VkFramebufferCreateInfo fbi{VK_STRUCTURE_TYPE_FRAMEBUFFER_CREATE_INFO, nullptr, 0, rp, 2, views, W, H, 1};
    VkFramebuffer fb;
    CHECK(vkCreateFramebuffer(dev, &fbi, nullptr, &fb));

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
    int32_t dyn[4] = {VK_DYNAMIC_STATE_VIEWPORT, VK_DYNAMIC_STATE_SCISSOR, VK_DYNAMIC_STATE_BLEND_CONSTANTS,
                      VK_DYNAMIC_STATE_DEPTH_BOUNDS};
    VkPipelineDynamicStateCreateInfo dys{VK_STRUCTURE_TYPE_PIPELINE_DYNAMIC_STATE_CREATE_INFO, nullptr, 0, 2, dyn};
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
    VkCommandBufferBeginInfo cbi{VK_STRUCTURE_TYPE_COMMAND_BUFFER_BEGIN_INFO, nullptr, 0, nullptr};
    CHECK(vkBeginCommandBuffer(cmd, &cbi));
    VkClearValue clears[2];
    for (int i = 0; i < 4; i++) clears[0].color.float32[i] = 0.2f;
    clears[1].depthStencil = VkClearDepthStencilValue{1.0f, 0};
    VkRenderPassBeginInfo rbi{VK_STRUCTURE_TYPE_RENDER_PASS_BEGIN_INFO, nullptr, rp, fb, {{0, 0}, {W, H}}, 2, clears};
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
    if (mode == "indirect") {
        const VkDrawIndirectCommand recs[2] = {{18, 1, 0, 0}, {18, 1, 18, 0}};
        std::vector<char> bytes((const char*)recs, (const char*)recs + sizeof(recs));
        buffer(bytes, &ibuf);
        vkCmdDrawIndirect(cmd, ibuf, 0, 2, sizeof(VkDrawIndirectCommand));
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
    CHECK(vkEndCommandBuffer(cmd));

    VkFenceCreateInfo fci{VK_STRUCTURE_TYPE_FENCE_CREATE_INFO, nullptr, 0};
    VkFence fence;
    CHECK(vkCreateFence(dev, &fci, nullptr, &fence));
    VkSubmitInfo si{VK_STRUCTURE_TYPE_SUBMIT_INFO, nullptr, 0, nullptr, nullptr, 1, &cmd, 0, nullptr};
    CHECK(vkQueueSubmit(queue, 1, &si, fence));
    CHECK(vkWaitForFences(dev, 1, &fence, VK_TRUE, ~0ull));

    std::ofstream(dir + "/color.bin", std::ios::binary).write((const char*)color_map, (std::streamsize)W * H * 4);
    std::ofstream(dir + "/depth.bin", std::ios::binary).write((const char*)depth_map, (std::streamsize)W * H * 2);
    vkDestroyDevice(dev, nullptr);
    vkDestroyInstance(g_inst, nullptr);
    std::printf("rendered %ux%u\n", W, H);
    return 0;
}
