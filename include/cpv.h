/*
 * cpv.h -- C-ABI of the MI355X draw path (libcpv.so).
 *
 * This is the thin layer between the host C++ Vulkan front end (the ICD,
 * libcpvulkan_amd_icd.so) and the HIP kernels that replace the reference's
 * CPU draw loop.  Every entry point takes plain pointers, sizes and POD
 * structs; no HIP or torch types appear here.
 *
 * What each entry point replaces in the reference (paths relative to the
 * reference repository):
 *   cpv_render_pass      -> BeginRenderPassCommand::Process clears
 *                           (CPVulkan/CommandBuffer.cpp:591-640) followed by
 *                           DrawCommand::Process / DrawIndexedCommand::Process
 *                           (CPVulkan/CommandBuffer.Draw.cpp:1777-1804, 1840-1864)
 *                           for every draw recorded inside the render pass.
 *   cpv_render_jobs      -> many independent render passes replayed by
 *                           Queue::Submit (CPVulkan/Queue.cpp:11-77), batched
 *                           into one persistent launch (no reference counterpart
 *                           for the batching; each job's result equals
 *                           cpv_render_pass on that job).
 *   cpv_program_create_spirv -> CompileVertexPipeline / CompileFragmentPipeline
 *                           (LLVMRuntime/PipelineCompiler.cpp:1801-1832), called
 *                           from GraphicsPipeline::Create (CPVulkan/Pipeline.cpp:599-714).
 *   cpv_alloc/cpv_free/cpv_upload/cpv_download -> DeviceMemory allocation and
 *                           vkMapMemory access (CPVulkan/Device.cpp:97-200);
 *                           cpv_alloc_host/cpv_free_host back host-visible
 *                           memory that vkMapMemory returns directly.
 *   cpv_device_copy      -> the transfer half of vkCmdCopyBuffer
 *                           (CPVulkan/CommandBuffer.Copy.cpp:15-60).
 *   cpv_copy_regions     -> the byte copies of CopyBufferCommand, CopyImageCommand,
 *                           CopyBufferToImageCommand and CopyImageToBufferCommand
 *                           (CPVulkan/CommandBuffer.Copy.cpp:38-1080), whose
 *                           region arithmetic the caller turns into rows.
 *   cpv_fill_buffer / cpv_update_buffer -> FillBufferCommand / UpdateBufferCommand
 *                           (CPVulkan/CommandBuffer.cpp:241-333).
 *   cpv_clear_rects      -> ClearImage for vkCmdClearColorImage and
 *                           vkCmdClearDepthStencilImage, and ClearAttachmentsCommand
 *                           (CPVulkan/CommandBuffer.Draw.cpp:117-149, 2032-2224, 2398-2419).
 *   cpv_blit_image       -> BlitImageCommand (CPVulkan/CommandBuffer.cpp:32-238).
 *   cpv_batch_*, cpv_autotune, cpv_mlp_*, cpv_set_predictor, cpv_get_tile_traces,
 *   cpv_set_config       -> no reference counterpart: recorded hipGraph
 *                           batches, the tile-configuration autotuner and the
 *                           learned tile scheduler only change how the same
 *                           frames are scheduled.
 * The Vulkan ICD entry points (vk_icdNegotiateLoaderICDInterfaceVersion,
 * vk_icdGetInstanceProcAddr, vk_icdGetPhysicalDeviceProcAddr; the reference's
 * CPVulkan/CPVulkan.cpp:14-104) are declared in cpv_vulkan.h and exported by
 * libcpvulkan_amd_icd.so, which is built on this C-ABI.
 *
 * Addresses in descriptor structs are device addresses obtained from
 * cpv_alloc (or any HIP device allocation).  The CPU oracle under oracle/
 * consumes the same structs with host addresses.
 *
 * Every function returns 0 on success or a negative CPV_ERROR_* code; the
 * message for the last failure on the calling thread is cpv_last_error().
 */
#ifndef CPV_H
#define CPV_H

#include <stddef.h>
#include <stdint.h>

#include "cpv_vk_enums.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CPV_MAX_VERTEX_BINDINGS 16
#define CPV_MAX_VERTEX_ATTRIBUTES 16
#define CPV_MAX_COLOR_ATTACHMENTS 1
#define CPV_MAX_UNIFORMS 4
#define CPV_MAX_TEXTURES 4
#define CPV_MAX_MIP_LEVELS 16

enum {
    CPV_OK = 0,
    CPV_ERROR_INVALID = -1,      /* malformed descriptor or argument */
    CPV_ERROR_UNSUPPORTED = -2,  /* state the reference aborts on, or not built here */
    CPV_ERROR_DEVICE = -3,       /* HIP runtime failure (VK_ERROR_DEVICE_LOST) */
    CPV_ERROR_OUT_OF_MEMORY = -4,
    CPV_ERROR_COMPILE = -5,      /* SPIR-V translation or hiprtc failure */
    CPV_ERROR_OVERFLOW = -6,     /* more (tile, primitive) pairs than a launch can hold (2^32 - 1, or the
                                    allocation-free entry point's array) */
    CPV_NOT_READY = 1,           /* cpv_fence_status: the work before the fence is still running */
    CPV_TIMEOUT = 2              /* cpv_fence_wait: the timeout passed first */
};

/* Built-in shader programs.  Each mirrors a GLSL shader of the benchmark
 * workloads; the SPIR-V path (cpv_program_create_spirv) compiles arbitrary
 * shaders of the supported subset into the same kernel templates. */
enum {
    CPV_VS_PASSTHROUGH = 1,   /* in vec4 pos @0, in vec4 color @1; gl_Position = pos; out vec4 = color */
    CPV_VS_MVP_COLOR = 2,     /* uniform(binding 0) mat4 mvp; gl_Position = mvp * pos; out vec4 = color
                                 (Samples/15-draw_cube/15-draw_cube.cpp:38-51) */
    CPV_VS_MVP_LAMBERT = 3,   /* in vec3 pos @0, vec3 normal @1, vec2 uv @2; uniform(binding 0) mat4 mvp;
                                 gl_Position = mvp * vec4(pos, 1); out vec3 normal; out vec2 uv */
    CPV_VS_PASSTHROUGH_POINT = 4  /* in vec4 pos @0, in vec4 color @1, in float size @2; gl_Position = pos;
                                     gl_PointSize = size; out vec4 = color (point topologies) */
};
enum {
    CPV_FS_COLOR = 1,         /* in vec4 color; out vec4 = color (15-draw_cube.cpp:53-61) */
    CPV_FS_TEX_LAMBERT = 2,   /* in vec3 n, in vec2 uv; sampler2D(binding 1); uniform(binding 2) vec4 L;
                                 out = texture(tex, uv) * max(dot(normalize(n), L.xyz), 0) */
    CPV_FS_COLOR_DISCARD = 3, /* in vec4 color; if (color.r < 0.25) discard; out vec4 = color */
    CPV_FS_COLOR_FLAT = 4,    /* flat in vec4 color; out vec4 = color (the provoking vertex's colour) */
    CPV_FS_COLOR_NOPERSPECTIVE = 5  /* noperspective in vec4 color; out vec4 = color */
};

/* Scheduling policies of the tile work queue (cpv_config.scheduler). */
enum {
    CPV_SCHED_GRID = 0,       /* one workgroup per tile, hardware dispatch order ("plain" path) */
    CPV_SCHED_COUNTER = 1,    /* persistent kernel, tiles taken in index order from an atomic counter */
    CPV_SCHED_PREDICTED = 2   /* hardware-dispatched tiles, the ones the learned cost predictor finds heaviest
                                 first (costliest first), the others in index order */
};

typedef struct cpv_context cpv_context;
typedef struct cpv_program cpv_program;

typedef struct cpv_vertex_binding_desc {
    uint32_t binding;
    uint32_t stride;
    uint32_t input_rate;      /* CPV_INPUT_RATE_* */
    uint32_t reserved;
} cpv_vertex_binding_desc;

typedef struct cpv_vertex_attribute_desc {
    uint32_t location;
    uint32_t binding;
    uint32_t format;          /* VkFormat */
    uint32_t offset;
} cpv_vertex_attribute_desc;

typedef struct cpv_stencil_state {
    uint32_t fail_op, pass_op, depth_fail_op, compare_op;
    uint32_t compare_mask, write_mask, reference, reserved;
} cpv_stencil_state;

typedef struct cpv_blend_attachment {
    uint32_t blend_enable;
    uint32_t src_color_factor, dst_color_factor, color_op;
    uint32_t src_alpha_factor, dst_alpha_factor, alpha_op;
    uint32_t color_write_mask;
} cpv_blend_attachment;

/* Everything GraphicsPipeline::Create freezes (CPVulkan/Pipeline.cpp:599-714,
 * CPVulkanBase/PipelineState.h:8-163), flattened. */
typedef struct cpv_pipeline_desc {
    uint32_t vs_kind;               /* CPV_VS_*, or 0 when program != NULL */
    uint32_t fs_kind;               /* CPV_FS_*, or 0 when program != NULL */
    cpv_program* program;           /* SPIR-V program from cpv_program_create_spirv */
    uint32_t topology;              /* CPV_TOPOLOGY_* */
    uint32_t primitive_restart;
    uint32_t binding_count;
    uint32_t attribute_count;
    cpv_vertex_binding_desc bindings[CPV_MAX_VERTEX_BINDINGS];
    cpv_vertex_attribute_desc attributes[CPV_MAX_VERTEX_ATTRIBUTES];
    uint32_t cull_mode, front_face, polygon_mode, rasterizer_discard;
    uint32_t depth_clamp, depth_test, depth_write, depth_compare;
    uint32_t depth_bounds_test, stencil_test, logic_op_enable, origin_upper_left;
    float min_depth_bounds, max_depth_bounds;
    cpv_stencil_state front, back;
    float blend_constants[4];
    uint32_t color_count;
    uint32_t color_formats[CPV_MAX_COLOR_ATTACHMENTS];
    uint32_t depth_stencil_format;  /* CPV_FORMAT_UNDEFINED when absent */
    cpv_blend_attachment blend[CPV_MAX_COLOR_ATTACHMENTS];
    float line_width;               /* RasterizationState.LineWidth (line topologies) */
    uint32_t reserved;
} cpv_pipeline_desc;

typedef struct cpv_sampler_desc {
    uint32_t mag_filter, min_filter, mipmap_mode;
    uint32_t address_u, address_v, address_w;
    float mip_lod_bias, min_lod, max_lod;
    uint32_t border_color;
    uint32_t unnormalized;
    uint32_t reserved;
} cpv_sampler_desc;

/* A combined image sampler: image base address (layer 0, level 0), the
 * linear tightly packed mip chain layout of GetNormalImageSize
 * (CPVulkanBase/Formats.cpp:455-482), and the sampler. */
typedef struct cpv_texture_desc {
    uint64_t address;
    uint32_t format, width, height, levels;
    uint32_t base_level, level_count;
    cpv_sampler_desc sampler;
} cpv_texture_desc;

typedef struct cpv_draw_desc {
    uint32_t indexed;
    uint32_t count;                 /* vertexCount or indexCount */
    uint32_t instance_count;
    uint32_t first;                 /* firstVertex or firstIndex */
    int32_t vertex_offset;
    uint32_t first_instance;
    uint32_t index_size;            /* 1, 2 or 4 bytes */
    uint32_t reserved;
    uint64_t index_address;         /* already advanced by the bound offset */
    uint64_t vertex_address[CPV_MAX_VERTEX_BINDINGS];   /* already advanced by bound offsets */
    uint64_t vertex_bytes[CPV_MAX_VERTEX_BINDINGS];     /* bytes readable from vertex_address */
    float viewport[6];              /* x, y, width, height, minDepth, maxDepth */
    uint32_t reserved2[2];
    uint64_t uniform_address[CPV_MAX_UNIFORMS];         /* uniform buffers of set 0 by binding */
    cpv_texture_desc textures[CPV_MAX_TEXTURES];        /* combined image samplers of set 0 by binding */
} cpv_draw_desc;

typedef struct cpv_attachment_desc {
    uint64_t address;               /* the image view's base level/layer */
    uint32_t format;
    uint32_t load_op;               /* CPV_LOAD_OP_* */
    union {
        float f32[4];
        uint32_t u32[4];
        int32_t i32[4];
        struct { float depth; uint32_t stencil; } ds;
    } clear;
} cpv_attachment_desc;

typedef struct cpv_pass_desc {
    uint32_t width, height;         /* attachment extent (all attachments share it) */
    uint32_t color_count;
    uint32_t has_depth_stencil;
    cpv_attachment_desc color[CPV_MAX_COLOR_ATTACHMENTS];
    cpv_attachment_desc depth_stencil;
} cpv_pass_desc;

/* One independent render job of a batch: a render pass with one draw. */
typedef struct cpv_job_desc {
    cpv_pass_desc pass;
    cpv_pipeline_desc pipeline;
    cpv_draw_desc draw;
} cpv_job_desc;

typedef struct cpv_config {
    uint32_t tile_size;             /* 8, 16 or 32 pixels */
    uint32_t waves_per_tile;        /* 1, 2 or 4 */
    uint32_t scheduler;             /* CPV_SCHED_* */
    uint32_t split_threshold;       /* tile list length above which a tile is split into chunks; 0 = never */
    uint32_t record_costs;          /* record per-tile cost traces (cpv_cost_trace) */
    uint32_t count_fragments;       /* count covered fragments exactly (cpv_last_fragment_count); this
                                       evaluates coverage for every pixel and so disables the exact
                                       hierarchical depth culling */
    uint32_t reserved[2];
} cpv_config;

/* Per-tile record of one rendered tile, written when cpv_config.record_costs
 * is set: the binning statistics that feed the cost predictor and the
 * measured cost (shader clock cycles spent on the tile). */
typedef struct cpv_tile_trace {
    float features[12];
    float cycles;
    uint32_t job, tile, reserved;
} cpv_tile_trace;

/* ---- context and memory ---- */
const char* cpv_last_error(void);
int cpv_context_create(int device_ordinal, cpv_context** out);
void cpv_context_destroy(cpv_context* ctx);
int cpv_alloc(cpv_context* ctx, uint64_t bytes, uint64_t* address);
int cpv_free(cpv_context* ctx, uint64_t address);
/* Pinned host memory the kernels read and write at the same address
 * (HOST_VISIBLE | HOST_COHERENT Vulkan memory); free with cpv_free_host. */
int cpv_alloc_host(cpv_context* ctx, uint64_t bytes, uint64_t* address);
int cpv_free_host(cpv_context* ctx, uint64_t address);
int cpv_upload(cpv_context* ctx, uint64_t dst, const void* src, uint64_t bytes);
int cpv_download(cpv_context* ctx, void* dst, uint64_t src, uint64_t bytes);
int cpv_synchronize(cpv_context* ctx);
/* Device-to-device copy on the context's stream (16-byte aligned addresses
 * and size); the transfer half of vkCmdCopyBuffer
 * (CPVulkan/CommandBuffer.Copy.cpp:15-60) for GPU-resident buffers. */
int cpv_device_copy(cpv_context* ctx, uint64_t dst, uint64_t src, uint64_t bytes);

/* ---------------------------------------------------------------- transfers
 * Each is enqueued on the context's stream, in order with render passes;
 * addresses are device addresses or host-visible (cpv_alloc_host) memory. */

/* rows x row_bytes bytes: row r goes from src + r * src_pitch to dst + r * dst_pitch. */
typedef struct cpv_copy_region {
    uint64_t src, dst;
    uint64_t src_pitch, dst_pitch;
    uint64_t row_bytes;
    uint32_t rows, pad;
} cpv_copy_region;

/* One mip level of one array layer of a linear image: rows of width x
 * pixel-size bytes with no padding (GetNormalImageSize, CPVulkanBase/Formats.cpp:455-482). */
typedef struct cpv_image_level {
    uint64_t address;
    uint32_t format, width, height, pad;
} cpv_image_level;

typedef struct cpv_rect {
    int32_t x, y;
    uint32_t width, height;
} cpv_rect;

/* VkImageBlit of one layer: corners (x0, y0, x1, y1) of the source and destination rectangles. */
typedef struct cpv_blit_region {
    int32_t src[4];
    int32_t dst[4];
    uint32_t aspect;   /* VkImageAspectFlags: COLOR 1, DEPTH 2, STENCIL 4 */
    uint32_t pad;
} cpv_blit_region;

int cpv_copy_regions(cpv_context* ctx, const cpv_copy_region* regions, uint32_t count);
/* vkCmdFillBuffer semantics of the reference, including its tail-byte placement for size % 4. */
int cpv_fill_buffer(cpv_context* ctx, uint64_t dst, uint64_t size, uint32_t data);
/* Copies `size` bytes of host data (staged at the call) to dst. */
int cpv_update_buffer(cpv_context* ctx, uint64_t dst, const void* data, uint64_t size);
/* Stores the packed clear value into every pixel of the rectangles.  value:
 * VkClearColorValue words for colour formats (float, int or uint by the
 * format's base type); depth float bits and stencil for depth/stencil formats. */
int cpv_clear_rects(cpv_context* ctx, const cpv_image_level* image, const uint32_t value[4],
                    const cpv_rect* rects, uint32_t count);
/* filter: VK_FILTER_NEAREST (0) or VK_FILTER_LINEAR (1); depth/stencil and integer formats take NEAREST. */
int cpv_blit_image(cpv_context* ctx, const cpv_image_level* src, const cpv_image_level* dst,
                   const cpv_blit_region* regions, uint32_t count, uint32_t filter);
int cpv_set_config(cpv_context* ctx, const cpv_config* config);
int cpv_get_config(cpv_context* ctx, cpv_config* config);

/* ---- rendering (asynchronous on the context's stream) ---- */
int cpv_render_pass(cpv_context* ctx, const cpv_pass_desc* pass,
                    const cpv_pipeline_desc* pipelines, const cpv_draw_desc* draws,
                    const uint32_t* draw_pipeline, uint32_t draw_count);
int cpv_render_jobs(cpv_context* ctx, const cpv_job_desc* jobs, uint32_t job_count);

/* cpv_render_pass without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory during a frame
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and asynchrony as cpv_render_pass;
 * every working buffer lives in one 64 MiB device array per GPU, set up
 * when the context is created and kept until the process ends.  No
 * hipMalloc, hipFree, hipHostMalloc or other allocating call is made.
 * A pass whose buffers (vertex outputs, triangle records, scans, tile
 * ranges) leave no room for pairs, or whose descriptors exceed 256 KiB, is
 * refused with CPV_ERROR_UNSUPPORTED before anything is enqueued.  A launch
 * whose (tile, primitive) pairs exceed the room left in the array draws
 * nothing, nor do the launches of the pass after it, and the next call on
 * the context (cpv_synchronize, for one) returns CPV_ERROR_OVERFLOW.
 * The CPV_SCHED_PREDICTED scheduler runs as CPV_SCHED_COUNTER here, and no
 * tile costs are recorded.  All calls on one GPU share its array, which
 * keeps the last call's contents: calls that overlap on one GPU (two
 * contexts, or two threads) race. */
int cpv_render_pass_noalloc(cpv_context* ctx, const cpv_pass_desc* pass,
                            const cpv_pipeline_desc* pipelines, const cpv_draw_desc* draws,
                            const uint32_t* draw_pipeline, uint32_t draw_count);

/* Several render passes in one call, each with its own draws: draws
 * first_draw[p] .. first_draw[p + 1] - 1 belong to passes[p] (first_draw has
 * pass_count + 1 entries, first_draw[0] = 0; draw_pipeline indexes
 * pipelines as in cpv_render_pass).  The result is that of one
 * cpv_render_pass per pass, in array order (a vkQueueSubmit replaying its
 * command buffers' render passes, CPVulkan/Queue.cpp:11-77).  Consecutive
 * passes that neither write one another's attachments nor sample them are
 * merged: the k-th run of draws sharing a shader program of every such pass
 * goes into one launch with one render target per pass, so N independent
 * single-program passes take one launch. */
int cpv_render_passes(cpv_context* ctx, const cpv_pass_desc* passes, uint32_t pass_count,
                      const cpv_pipeline_desc* pipelines, const cpv_draw_desc* draws,
                      const uint32_t* draw_pipeline, const uint32_t* first_draw);

/* Raster launches enqueued on the context so far (render calls and batch
 * submits; the launches of a batch submit count once per submit). */
int cpv_launch_count(cpv_context* ctx, uint64_t* launches);

/* Number of (primitive, pixel) pairs that passed coverage and culling in the
 * last cpv_render_pass / cpv_render_jobs call (synchronises). */
int cpv_last_fragment_count(cpv_context* ctx, uint64_t* fragments);

/* Device-side duration in milliseconds of the last render call's raster
 * kernel and of the whole call, measured with HIP events on the context's
 * stream (synchronises). */
int cpv_last_timing(cpv_context* ctx, float* raster_ms, float* total_ms);

/* Copies up to capacity tile traces of the last render call with
 * record_costs set; *count receives the number available. */
int cpv_get_tile_traces(cpv_context* ctx, cpv_tile_trace* traces, uint32_t capacity, uint32_t* count);

/* ---- recorded batches ----
 * A batch is a set of render jobs recorded once and replayed many times, the
 * way a Vulkan command buffer is recorded by vkCmdDraw* and replayed by every
 * vkQueueSubmit (CPVulkan/CommandBuffer.Draw.cpp:2350-2360, Queue.cpp:52-74).
 * Recording plans every launch, uploads the descriptors, runs the vertex
 * and setup stages once to size the (tile, primitive) pair arrays, and
 * captures every kernel of the frame -- all jobs, one launch per shader
 * program -- into hipGraphs; it uses the context's configuration at
 * recording time (tile size, waves per tile, scheduler, predictor).  A
 * submit replays the graphs on the context's streams without host work.
 * Consecutive submits of one batch with nothing else enqueued on the context
 * in between are pipelined: a submit's vertex, setup and binning stages
 * (which touch no attachment) run while the previous submit's raster kernel
 * does, on two alternating sets of working buffers; the results are those
 * of running the submits one after the other.
 * Vertex and index data may change between submits (not while a submit is in
 * flight).  A submit that needs more (tile, primitive) pairs than were
 * recorded draws nothing; the next call on the context finds it, records
 * the batch again with the pair counts it needed and replays that submit
 * before anything else is enqueued, so the results are those of the
 * submits run in order.  CPV_ERROR_OVERFLOW is returned only for more than
 * 2^32 - 1 pairs in one launch.  With the learned scheduler the tile order
 * is computed once, at recording time, from the geometry then bound, and
 * every submit reuses it (a schedule only: the frames are exact whatever
 * the geometry of later submits). */
typedef struct cpv_batch cpv_batch;
int cpv_batch_create(cpv_context* ctx, const cpv_job_desc* jobs, uint32_t job_count, cpv_batch** out);
/* A batch of render passes laid out as for cpv_render_passes (a command
 * buffer's render passes, recorded once). */
int cpv_batch_create_passes(cpv_context* ctx, const cpv_pass_desc* passes, uint32_t pass_count,
                            const cpv_pipeline_desc* pipelines, const cpv_draw_desc* draws,
                            const uint32_t* draw_pipeline, const uint32_t* first_draw, cpv_batch** out);
int cpv_batch_submit(cpv_batch* batch);
/* Waits for the last submit; *fragments (optional) receives its covered
 * fragment count when the batch was recorded with count_fragments on. */
int cpv_batch_wait(cpv_batch* batch, uint64_t* fragments);
/* Device time of the last submit in milliseconds (HIP events around the graph). */
int cpv_batch_timing(cpv_batch* batch, float* ms);
void cpv_batch_destroy(cpv_batch* batch);

/* ---- fences ----
 * A fence marks the end of the work enqueued on its context up to
 * cpv_fence_signal (a VkFence signalled by vkQueueSubmit,
 * CPVulkan/Queue.cpp:68-71).  cpv_fence_status returns CPV_OK once that work
 * has finished (or when the fence was created signalled and not signalled
 * or reset since) and CPV_NOT_READY before; cpv_fence_wait returns CPV_OK or
 * CPV_TIMEOUT after timeout_ns nanoseconds (UINT64_MAX: no limit), as
 * Fence::Wait and getStatus do (CPVulkan/Fence.cpp:47-56, 157-160). */
typedef struct cpv_fence cpv_fence;
int cpv_fence_create(cpv_context* ctx, int signaled, cpv_fence** out);
int cpv_fence_signal(cpv_fence* fence);
int cpv_fence_reset(cpv_fence* fence);
int cpv_fence_status(cpv_fence* fence);
int cpv_fence_wait(cpv_fence* fence, uint64_t timeout_ns);
void cpv_fence_destroy(cpv_fence* fence);

/* ---- autotuner ----
 * Records the jobs as a batch under every tile configuration (tile size 8,
 * 16, 32 x waves per tile 1, 2, 4 where the tile has that many 8x8 blocks),
 * replays each `repeats` times after one warm-up replay, and returns the
 * mean device time per configuration and the fastest configuration (the
 * context's configuration with tile_size / waves_per_tile replaced).  The
 * context's configuration is left unchanged. */
typedef struct cpv_autotune_entry {
    uint32_t tile_size;
    uint32_t waves_per_tile;
    float ms;
    uint32_t reserved;
} cpv_autotune_entry;
int cpv_autotune(cpv_context* ctx, const cpv_job_desc* jobs, uint32_t job_count, uint32_t repeats,
                 cpv_autotune_entry* entries, uint32_t capacity, uint32_t* count, cpv_config* best);

/* ---- tile-cost predictor (CPV_SCHED_PREDICTED) ----
 * A BF16 MLP (12 tile features -> 64 -> 64 -> 1, ReLU, FP32 accumulate) that
 * predicts log2 of a tile's raster cycles from its binning statistics.  It
 * is trained on the device by hand-written MFMA forward and backward
 * kernels on cost traces the raster kernel records (cpv_get_tile_traces).
 * The persistent raster kernel takes tiles longest-predicted first; the
 * prediction changes only the order in which tiles are taken, never a
 * pixel.
 *
 * Parameters, fp32, in this order (CPV_MLP_PARAMS floats):
 *   W1[12][64], b1[64], W2[64][64], b2[64], w3[64], b3
 * with h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2), y = h2 . w3 + b3. */
#define CPV_MLP_FEATURES 12
#define CPV_MLP_HIDDEN 64
#define CPV_MLP_PARAMS (12 * 64 + 64 + 64 * 64 + 64 + 64 + 1)
typedef struct cpv_mlp cpv_mlp;
int cpv_mlp_create(cpv_context* ctx, uint64_t seed, cpv_mlp** out);
void cpv_mlp_destroy(cpv_mlp* mlp);
int cpv_mlp_get_params(cpv_mlp* mlp, float* params, uint32_t count);
int cpv_mlp_set_params(cpv_mlp* mlp, const float* params, uint32_t count);
/* y[rows] = network(x[rows][12]) (host arrays). */
int cpv_mlp_predict(cpv_mlp* mlp, const float* x, uint32_t rows, float* y);
/* One forward and backward pass over rows (host arrays): the gradient of
 * the mean squared error mean((y - target)^2) in the parameter order above,
 * and the loss. */
int cpv_mlp_gradients(cpv_mlp* mlp, const float* x, const float* target, uint32_t rows, float* grads, float* loss);
/* Adam training on the device: `steps` minibatches of `batch` rows drawn
 * with a seeded generator from (x, target); *loss receives the mean squared
 * error over all rows after training. */
int cpv_mlp_train(cpv_mlp* mlp, const float* x, const float* target, uint32_t rows, uint32_t steps, uint32_t batch,
                  float learning_rate, uint64_t seed, float* loss);
/* The predictor CPV_SCHED_PREDICTED uses (NULL: the scheduler falls back to
 * the counter-ordered queue). */
int cpv_set_predictor(cpv_context* ctx, cpv_mlp* mlp);

/* ---- shader programs ---- */
int cpv_program_create_spirv(cpv_context* ctx,
                             const uint32_t* vs_words, size_t vs_word_count, const char* vs_entry,
                             const uint32_t* fs_words, size_t fs_word_count, const char* fs_entry,
                             cpv_program** out);
void cpv_program_destroy(cpv_context* ctx, cpv_program* program);
/* The HIP source a (vertex, fragment) SPIR-V pair translates to, without
 * compiling it (no device needed).  Writes at most capacity bytes including
 * the terminating NUL; *length receives the full length.  On
 * CPV_ERROR_UNSUPPORTED, cpv_spirv_translate_error() names the construct. */
int cpv_spirv_translate(const uint32_t* vs_words, size_t vs_word_count, const char* vs_entry,
                        const uint32_t* fs_words, size_t fs_word_count, const char* fs_entry,
                        char* source, size_t capacity, size_t* length);
const char* cpv_spirv_translate_error(void);
/* Translates and compiles the pair with hiprtc for gfx950 without loading
 * the code object (no device needed); *code_size receives its size. */
int cpv_spirv_build(const uint32_t* vs_words, size_t vs_word_count, const char* vs_entry,
                    const uint32_t* fs_words, size_t fs_word_count, const char* fs_entry, size_t* code_size);
/* Generated HIP source of a program (for inspection); valid until destroy. */
const char* cpv_program_source(cpv_program* program);

/* sizeof() of the descriptor structs, in the order: pipeline, draw, pass,
 * attachment, texture, sampler, job, config, tile trace, autotune entry.  Lets bindings in
 * other languages check their struct layouts.  Returns the count. */
int cpv_abi_sizes(uint32_t* out, uint32_t capacity);

#ifdef __cplusplus
}
#endif

#endif
