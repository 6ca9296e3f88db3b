/*
 * cpvref -- CPU oracle for the CPVulkan draw path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, and by nothing else.  The product (libcpv.so,
 * the ICD) never links, loads or calls it.
 *
 * It is a single-threaded C restatement of the reference's draw loop and
 * its JIT'd per-pipeline / per-format code, consuming the same descriptor
 * structs as the C-ABI in include/cpv.h but with HOST addresses.  Each
 * function cites the reference file:line it follows (paths relative to the
 * reference repository).
 */
#ifndef CPVREF_H
#define CPVREF_H

#include <stdint.h>

#include "../include/cpv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Whole render pass: clears per load_op, then every draw in order.
 * *fragments receives the number of (primitive, pixel) pairs that passed
 * coverage and culling.  Returns 0 or a negative CPV_ERROR_* code. */
int cpvref_render_pass(const cpv_pass_desc* pass, const cpv_pipeline_desc* pipelines,
                       const cpv_draw_desc* draws, const uint32_t* draw_pipeline,
                       uint32_t draw_count, uint64_t* fragments);

/* Same, restricted to primitives [prim_begin, prim_end) of each draw
 * (used to time a bounded sample of a large draw). */
int cpvref_render_pass_range(const cpv_pass_desc* pass, const cpv_pipeline_desc* pipelines,
                             const cpv_draw_desc* draws, const uint32_t* draw_pipeline,
                             uint32_t draw_count, uint64_t prim_begin, uint64_t prim_end,
                             uint64_t* fragments);

/* Same result as cpvref_render_pass, bit for bit, computed tile-parallel
 * with OpenMP over `threads` host threads: the framebuffer is cut into
 * 64x64 tiles, the triangles are set up once and binned to the tiles in
 * primitive order, and every tile runs the reference's per-pixel loop over
 * its own pixels (threads <= 1: the single-threaded replay). */
int cpvref_render_pass_parallel(const cpv_pass_desc* pass, const cpv_pipeline_desc* pipelines,
                                const cpv_draw_desc* draws, const uint32_t* draw_pipeline,
                                uint32_t draw_count, int threads, uint64_t* fragments);

/* Format conversions (exposed for the format tests). */
int cpvref_pack_f32(uint32_t format, const float rgba[4], uint8_t* dst);
int cpvref_unpack_f32(uint32_t format, const uint8_t* src, float rgba[4]);
int cpvref_pack_u32(uint32_t format, const uint32_t rgba[4], uint8_t* dst);
int cpvref_pack_i32(uint32_t format, const int32_t rgba[4], uint8_t* dst);
int cpvref_pack_depth_stencil(uint32_t format, float depth, uint32_t stencil, uint8_t* dst);
float cpvref_unpack_depth(uint32_t format, const uint8_t* src);
uint32_t cpvref_unpack_stencil(uint32_t format, const uint8_t* src);
uint32_t cpvref_format_size(uint32_t format);

/* Mip chain layout of GetNormalImageSize: level offsets (bytes) and total. */
uint64_t cpvref_image_layout(uint32_t format, uint32_t width, uint32_t height, uint32_t levels,
                             uint64_t* level_offsets);

/* Texture sampling through ImageSampleImplicitLod (lambda = 0),
 * ImageSampleExplicitLod (lambda base = lod) and ImageFetch. */
int cpvref_sample_implicit(const cpv_texture_desc* texture, float u, float v, float rgba[4]);
int cpvref_sample_explicit(const cpv_texture_desc* texture, float u, float v, float lod, float rgba[4]);
int cpvref_fetch(const cpv_texture_desc* texture, int32_t x, int32_t y, float rgba[4]);

/* Transfers, with host addresses in the same structs as the C-ABI. */
void cpvref_fill_buffer(uint8_t* dst, uint64_t size, uint32_t data);
int cpvref_clear_rects(const cpv_image_level* image, const uint32_t value[4], const cpv_rect* rects, uint32_t count);
int cpvref_blit_image(const cpv_image_level* src, const cpv_image_level* dst, const cpv_blit_region* regions, uint32_t count,
                      uint32_t filter);

/* Blend of one fragment (ApplyBlend, dead code in the reference). */
void cpvref_apply_blend(const float src[4], const float dst[4], const float constant[4],
                        const cpv_blend_attachment* state, float out[4]);

/* The fragment wrapper after the shader (PipelineCompiler.cpp:1087-1727:
 * depth bounds, stencil, depth test and writes, colour write with blend)
 * for one fragment of colour `colour` at pixel (x, y) of the pass's
 * attachments, with the viewport-transformed depth `depth`. */
int cpvref_fragment(const cpv_pipeline_desc* pipeline, const cpv_pass_desc* pass, int32_t x, int32_t y, float depth,
                    int front, const float colour[4]);

/* The shaders' vector math in glm 0.9.5.3's operation order: op 0
 * out = a * b (a a column-major mat4, b a vec4), op 1 out[0] = dot(a, b)
 * (vec3), op 2 out = normalize(a) (vec3). */
void cpvref_glm(uint32_t op, const float* a, const float* b, float* out);

/* One edge function value and the coverage decision for one pixel of one
 * triangle exactly as ProcessTriangles/GetFragmentInput compute them.
 * positions: 3 clip-space vec4 (post-VS, before any swap). Returns 1 if the
 * pixel is covered (and not culled); fills depth, front, weights. */
int cpvref_pixel_coverage(const float positions[12], uint32_t cull_mode, uint32_t front_face,
                          float vp_width, float vp_height, int32_t x, int32_t y,
                          float* depth, int* front, float weights[3]);

#ifdef __cplusplus
}
#endif

#endif
