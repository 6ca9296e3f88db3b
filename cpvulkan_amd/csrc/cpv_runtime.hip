// Host runtime of libcpv.so, part 1: errors, the format table, context
// and memory entry points, configuration, the shader-program kernel tables
// and the translation of C-ABI descriptors into the kernels' layouts.  The
// launch sequence lives in cpv_launch.hip, the tile-cost predictor in
// cpv_mlp.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cpv_internal.h"
#include "programs_builtin.h"
#include "raster_kernels.h"

using namespace cpv;
using namespace cpvrt;

// ------------------------------------------------------------------ errors

static thread_local std::string g_error;

int cpvrt::fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

extern "C" const char* cpv_last_error(void) { return g_error.c_str(); }

// ----------------------------------------------------------- format table

// Normal and depth/stencil rows of the reference's format table
// (CPVulkanBase/Formats.cpp:209-444) for the formats the draw path handles.
bool cpvrt::format_desc(uint32_t f, FormatDesc* o)
{
    std::memset(o, 0, sizeof(*o));
    o->vk_format = f;
    for (int i = 0; i < 4; i++) o->off[i] = -1;
    o->depth_off = o->stencil_off = -1;
    int ch = 0, el = 1, base = kBaseNone;
    bool bgr = false;
    switch (f) {
    case CPV_FORMAT_R8_UNORM: ch = 1; base = kBaseUNorm; break;
    case CPV_FORMAT_R8_SNORM: ch = 1; base = kBaseSNorm; break;
    case CPV_FORMAT_R8_USCALED: ch = 1; base = kBaseUScaled; break;
    case CPV_FORMAT_R8_SSCALED: ch = 1; base = kBaseSScaled; break;
    case CPV_FORMAT_R8_UINT: ch = 1; base = kBaseUInt; break;
    case CPV_FORMAT_R8_SINT: ch = 1; base = kBaseSInt; break;
    case CPV_FORMAT_R8G8_UNORM: ch = 2; base = kBaseUNorm; break;
    case CPV_FORMAT_R8G8_SNORM: ch = 2; base = kBaseSNorm; break;
    case CPV_FORMAT_R8G8_UINT: ch = 2; base = kBaseUInt; break;
    case CPV_FORMAT_R8G8_SINT: ch = 2; base = kBaseSInt; break;
    case CPV_FORMAT_R8G8B8_UNORM: ch = 3; base = kBaseUNorm; break;
    case CPV_FORMAT_R8G8B8_SNORM: ch = 3; base = kBaseSNorm; break;
    case CPV_FORMAT_R8G8B8_UINT: ch = 3; base = kBaseUInt; break;
    case CPV_FORMAT_R8G8B8_SINT: ch = 3; base = kBaseSInt; break;
    case CPV_FORMAT_B8G8R8_UNORM: ch = 3; base = kBaseUNorm; bgr = true; break;
    case CPV_FORMAT_B8G8R8_SNORM: ch = 3; base = kBaseSNorm; bgr = true; break;
    case CPV_FORMAT_B8G8R8_UINT: ch = 3; base = kBaseUInt; bgr = true; break;
    case CPV_FORMAT_B8G8R8_SINT: ch = 3; base = kBaseSInt; bgr = true; break;
    case CPV_FORMAT_R8G8B8A8_UNORM: ch = 4; base = kBaseUNorm; break;
    case CPV_FORMAT_R8G8B8A8_SNORM: ch = 4; base = kBaseSNorm; break;
    case CPV_FORMAT_R8G8B8A8_USCALED: ch = 4; base = kBaseUScaled; break;
    case CPV_FORMAT_R8G8B8A8_SSCALED: ch = 4; base = kBaseSScaled; break;
    case CPV_FORMAT_R8G8B8A8_UINT: ch = 4; base = kBaseUInt; break;
    case CPV_FORMAT_R8G8B8A8_SINT: ch = 4; base = kBaseSInt; break;
    case CPV_FORMAT_B8G8R8A8_UNORM: ch = 4; base = kBaseUNorm; bgr = true; break;
    case CPV_FORMAT_B8G8R8A8_SNORM: ch = 4; base = kBaseSNorm; bgr = true; break;
    case CPV_FORMAT_B8G8R8A8_UINT: ch = 4; base = kBaseUInt; bgr = true; break;
    case CPV_FORMAT_B8G8R8A8_SINT: ch = 4; base = kBaseSInt; bgr = true; break;
    case CPV_FORMAT_R16_UNORM: ch = 1; el = 2; base = kBaseUNorm; break;
    case CPV_FORMAT_R16_SNORM: ch = 1; el = 2; base = kBaseSNorm; break;
    case CPV_FORMAT_R16_UINT: ch = 1; el = 2; base = kBaseUInt; break;
    case CPV_FORMAT_R16_SINT: ch = 1; el = 2; base = kBaseSInt; break;
    case CPV_FORMAT_R16G16_UNORM: ch = 2; el = 2; base = kBaseUNorm; break;
    case CPV_FORMAT_R16G16_UINT: ch = 2; el = 2; base = kBaseUInt; break;
    case CPV_FORMAT_R16G16_SINT: ch = 2; el = 2; base = kBaseSInt; break;
    case CPV_FORMAT_R16G16B16A16_UNORM: ch = 4; el = 2; base = kBaseUNorm; break;
    case CPV_FORMAT_R16G16B16A16_SNORM: ch = 4; el = 2; base = kBaseSNorm; break;
    case CPV_FORMAT_R16G16B16A16_UINT: ch = 4; el = 2; base = kBaseUInt; break;
    case CPV_FORMAT_R16G16B16A16_SINT: ch = 4; el = 2; base = kBaseSInt; break;
    case CPV_FORMAT_R32_UINT: ch = 1; el = 4; base = kBaseUInt; break;
    case CPV_FORMAT_R32_SINT: ch = 1; el = 4; base = kBaseSInt; break;
    case CPV_FORMAT_R32_SFLOAT: ch = 1; el = 4; base = kBaseSFloat; break;
    case CPV_FORMAT_R32G32_UINT: ch = 2; el = 4; base = kBaseUInt; break;
    case CPV_FORMAT_R32G32_SINT: ch = 2; el = 4; base = kBaseSInt; break;
    case CPV_FORMAT_R32G32_SFLOAT: ch = 2; el = 4; base = kBaseSFloat; break;
    case CPV_FORMAT_R32G32B32_UINT: ch = 3; el = 4; base = kBaseUInt; break;
    case CPV_FORMAT_R32G32B32_SINT: ch = 3; el = 4; base = kBaseSInt; break;
    case CPV_FORMAT_R32G32B32_SFLOAT: ch = 3; el = 4; base = kBaseSFloat; break;
    case CPV_FORMAT_R32G32B32A32_UINT: ch = 4; el = 4; base = kBaseUInt; break;
    case CPV_FORMAT_R32G32B32A32_SINT: ch = 4; el = 4; base = kBaseSInt; break;
    case CPV_FORMAT_R32G32B32A32_SFLOAT: ch = 4; el = 4; base = kBaseSFloat; break;
    case CPV_FORMAT_D16_UNORM: o->total = 2; o->element = 2; o->base = kBaseUNorm; o->depth_off = 0; o->depth_kind = 16; return true;
    case CPV_FORMAT_X8_D24_UNORM_PACK32: o->total = 4; o->element = 4; o->base = kBaseUNorm; o->depth_off = 0; o->depth_kind = 24; return true;
    case CPV_FORMAT_D32_SFLOAT: o->total = 4; o->element = 4; o->base = kBaseSFloat; o->depth_off = 0; o->depth_kind = 32; return true;
    case CPV_FORMAT_S8_UINT: o->total = 1; o->element = 1; o->base = kBaseUInt; o->stencil_off = 0; return true;
    case CPV_FORMAT_D16_UNORM_S8_UINT: o->total = 3; o->element = 2; o->base = kBaseUNorm; o->depth_off = 0; o->stencil_off = 2; o->depth_kind = 16; return true;
    case CPV_FORMAT_D24_UNORM_S8_UINT: o->total = 4; o->element = 3; o->base = kBaseUNorm; o->depth_off = 0; o->stencil_off = 3; o->depth_kind = 24; return true;
    case CPV_FORMAT_D32_SFLOAT_S8_UINT: o->total = 8; o->element = 4; o->base = kBaseSFloat; o->depth_off = 0; o->stencil_off = 4; o->depth_kind = 32; return true;
    default: return false;
    }
    o->base = base;
    o->element = el;
    o->total = el * ch;
    static const int rgba[4] = {0, 1, 2, 3}, bgra[4] = {2, 1, 0, 3};
    for (int i = 0; i < ch; i++) o->off[i] = (bgr ? bgra[i] : rgba[i]) * el;
    return true;
}

static bool is_depth_stencil(const FormatDesc& f) { return f.depth_off >= 0 || f.stencil_off >= 0; }

// ------------------------------------------------------------- utilities

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

void cpvrt::fill_u32(hipStream_t s, void* p, uint32_t v, uint64_t n)
{
    if (!n) return;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, s, (uint32_t*)p, v, n);
}

int DevBuf::ensure(size_t bytes, bool frozen)
{
    if (bytes <= cap) return CPV_OK;
    if (frozen) return fail(CPV_ERROR_INVALID, "buffer of %zu bytes needs %zu while a launch is recorded", cap, bytes);
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) return fail(CPV_ERROR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
    cap = want;
    return CPV_OK;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

void Buffers::release_all()
{
    DevBuf* bufs[] = {&vout, &recs, &counts, &offsets, &keys[0], &keys[1], &vals[0], &vals[1], &tile_start, &hist,
                      &hist_base, &sums, &big, &arena, &stats, &traces, &features, &pred, &okeys[0], &okeys[1],
                      &ovals[0], &ovals[1], &line_ext};
    for (DevBuf* b : bufs) b->release();
}

int cpvrt::stage_upload(cpv_context* c, void* dst, const void* src, size_t bytes, bool may_grow, hipStream_t s)
{
    if (!bytes) return CPV_OK;
    cpv_context::Staged& slot = c->staged[c->staged_next++ % 4u];
    if (!slot.done) HIP_TRY(hipEventCreateWithFlags(&slot.done, hipEventDisableTiming));
    if (slot.used) HIP_TRY(hipEventSynchronize(slot.done));   // its last copy has read it
    if (bytes > slot.cap && !may_grow)
        return fail(CPV_ERROR_UNSUPPORTED, "%zu bytes of descriptors exceed the %zu-byte staging slot", bytes, slot.cap);
    if (bytes > slot.cap) {
        if (slot.p) (void)hipHostFree(slot.p);
        slot.p = nullptr;
        slot.cap = 0;
        const size_t want = std::max<size_t>(bytes * 2, 64 << 10);
        HIP_TRY(hipHostMalloc(&slot.p, want, hipHostMallocDefault));
        slot.cap = want;
    }
    std::memcpy(slot.p, src, bytes);
    if (!s) s = c->stream;
    HIP_TRY(hipMemcpyAsync(dst, slot.p, bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(slot.done, s));
    slot.used = true;
    return CPV_OK;
}

extern "C" int cpv_context_create(int device_ordinal, cpv_context** out)
{
    if (!out) return fail(CPV_ERROR_INVALID, "out is null");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device_ordinal < 0 || device_ordinal >= n) return fail(CPV_ERROR_INVALID, "device %d of %d", device_ordinal, n);
    HIP_TRY(hipSetDevice(device_ordinal));
    cpv_context* c = new cpv_context();
    c->device = device_ordinal;
    c->cfg.tile_size = 16;
    c->cfg.waves_per_tile = 4;
    c->cfg.scheduler = CPV_SCHED_GRID;
    c->cfg.count_fragments = 1;
    if (hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, device_ordinal) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev_begin) != hipSuccess || hipEventCreate(&c->ev_end) != hipSuccess ||
        hipEventCreate(&c->ev_raster0) != hipSuccess || hipEventCreate(&c->ev_raster1) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_check, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_pre, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->set_front_done[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->set_front_done[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->set_raster_done[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->set_raster_done[1], hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&c->front_stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&c->readback, 256, hipHostMallocDefault) != hipSuccess) {
        delete c;
        return fail(CPV_ERROR_DEVICE, "stream/event/pinned allocation failed");
    }
    if (c->buf.stats.ensure(kStatAlloc * 4)) { delete c; return CPV_ERROR_OUT_OF_MEMORY; }
    c->last_stats = c->buf.stats.as<uint32_t>();
    // the staging slots start large enough for the allocation-free entry point
    for (cpv_context::Staged& slot : c->staged) {
        if (hipEventCreateWithFlags(&slot.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc(&slot.p, kStagingSlotBytes, hipHostMallocDefault) != hipSuccess) {
            cpv_context_destroy(c);
            return fail(CPV_ERROR_DEVICE, "staging slot allocation failed");
        }
        slot.cap = kStagingSlotBytes;
    }
    // the scratch array's address on this device; a failure here is reported
    // by cpv_render_pass_noalloc, the only user of the array
    if (scratch_setup(c) != CPV_OK) c->scratch = nullptr;
    *out = c;
    return CPV_OK;
}

extern "C" void cpv_context_destroy(cpv_context* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)sync_context(c);
    destroy_pending(c->pending);
    c->pending = nullptr;
    c->buf.release_all();
    for (cpvrt::Buffers& b : c->lbuf) b.release_all();
    for (hipEvent_t e : {c->ev_pre, c->set_front_done[0], c->set_front_done[1], c->set_raster_done[0], c->set_raster_done[1]})
        if (e) (void)hipEventDestroy(e);
    for (cpv_context::Staged& slot : c->staged) {
        if (slot.p) (void)hipHostFree(slot.p);
        if (slot.done) (void)hipEventDestroy(slot.done);
    }
    if (c->readback) (void)hipHostFree(c->readback);
    (void)hipEventDestroy(c->ev_begin);
    (void)hipEventDestroy(c->ev_end);
    (void)hipEventDestroy(c->ev_raster0);
    (void)hipEventDestroy(c->ev_raster1);
    (void)hipEventDestroy(c->ev_check);
    (void)hipStreamDestroy(c->stream);
    if (c->front_stream) {
        (void)hipStreamSynchronize(c->front_stream);
        (void)hipStreamDestroy(c->front_stream);
    }
    delete c;
}


extern "C" int cpv_alloc(cpv_context* c, uint64_t bytes, uint64_t* address)
{
    if (!c || !address) return fail(CPV_ERROR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(CPV_ERROR_OUT_OF_MEMORY, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
    *address = (uint64_t)(uintptr_t)p;
    return CPV_OK;
}

extern "C" int cpv_free(cpv_context* c, uint64_t address)
{
    if (!c) return fail(CPV_ERROR_INVALID, "null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree((void*)(uintptr_t)address));
    return CPV_OK;
}

extern "C" int cpv_alloc_host(cpv_context* c, uint64_t bytes, uint64_t* address)
{
    if (!c || !address) return fail(CPV_ERROR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) return fail(CPV_ERROR_OUT_OF_MEMORY, "hipHostMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
    *address = (uint64_t)(uintptr_t)p;
    return CPV_OK;
}

extern "C" int cpv_free_host(cpv_context* c, uint64_t address)
{
    if (!c) return fail(CPV_ERROR_INVALID, "null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipHostFree((void*)(uintptr_t)address));
    return CPV_OK;
}

extern "C" int cpv_upload(cpv_context* c, uint64_t dst, const void* src, uint64_t bytes)
{
    if (!c || (!src && bytes)) return fail(CPV_ERROR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    note_enqueue(c);
    if (int rc = resolve_pending(c)) return rc;
    HIP_TRY(hipMemcpyAsync((void*)(uintptr_t)dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CPV_OK;
}

extern "C" int cpv_download(cpv_context* c, void* dst, uint64_t src, uint64_t bytes)
{
    if (!c || (!dst && bytes)) return fail(CPV_ERROR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    note_enqueue(c);
    if (int rc = resolve_pending(c)) return rc;
    HIP_TRY(hipMemcpyAsync(dst, (const void*)(uintptr_t)src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CPV_OK;
}

__global__ void copy_u32x4_kernel(uint4* dst, const uint4* src, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

extern "C" int cpv_device_copy(cpv_context* c, uint64_t dst, uint64_t src, uint64_t bytes)
{
    if (!c) return fail(CPV_ERROR_INVALID, "null context");
    if ((dst | src | bytes) & 15u) return fail(CPV_ERROR_INVALID, "device copy needs 16-byte aligned addresses and size");
    HIP_TRY(hipSetDevice(c->device));
    note_enqueue(c);
    if (int rc = resolve_pending(c)) return rc;
    const uint64_t n = bytes / 16;
    if (n) {
        hipLaunchKernelGGL(copy_u32x4_kernel, dim3((uint32_t)std::min<uint64_t>((n + 255) / 256, 65536)), dim3(256), 0, c->stream,
                           (uint4*)(uintptr_t)dst, (const uint4*)(uintptr_t)src, n);
        HIP_TRY(hipGetLastError());
    }
    return CPV_OK;
}

extern "C" int cpv_synchronize(cpv_context* c)
{
    if (!c) return fail(CPV_ERROR_INVALID, "null context");
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = resolve_pending(c)) return rc;
    HIP_TRY(sync_context(c));
    return CPV_OK;
}

bool cpvrt::valid_tile_config(uint32_t t, uint32_t w)
{
    if (t == 8) return w == 1;
    if (t == 16 || t == 32) return w == 1 || w == 2 || w == 4;
    return false;
}

extern "C" int cpv_set_config(cpv_context* c, const cpv_config* cfg)
{
    if (!c || !cfg) return fail(CPV_ERROR_INVALID, "null argument");
    if (!valid_tile_config(cfg->tile_size, cfg->waves_per_tile))
        return fail(CPV_ERROR_INVALID, "unsupported tile %u x waves %u", cfg->tile_size, cfg->waves_per_tile);
    if (cfg->scheduler > CPV_SCHED_PREDICTED) return fail(CPV_ERROR_INVALID, "scheduler %u", cfg->scheduler);
    c->cfg = *cfg;
    return CPV_OK;
}

extern "C" int cpv_get_config(cpv_context* c, cpv_config* cfg)
{
    if (!c || !cfg) return fail(CPV_ERROR_INVALID, "null argument");
    *cfg = c->cfg;
    return CPV_OK;
}

// --------------------------------------------------------- kernel tables

namespace cpvrt {

using RasterFn = void (*)(RasterArgs);
using VertexFn = void (*)(VertexArgs);

int tile_config_index(uint32_t t, uint32_t w)
{
    if (t == 8) return 0;
    if (t == 16) return w == 1 ? 1 : w == 2 ? 2 : 3;
    return w == 1 ? 4 : w == 2 ? 5 : 6;
}

bool builtin_program(uint32_t vs, uint32_t fs, Program* p)
{
    switch (vs) {
    case CPV_VS_PASSTHROUGH: p->vertex.host_fn = (const void*)&vertex_kernel<VsPassthrough>; p->vs_outputs = VsPassthrough::kOutputs; break;
    case CPV_VS_MVP_COLOR: p->vertex.host_fn = (const void*)&vertex_kernel<VsMvpColor>; p->vs_outputs = VsMvpColor::kOutputs; break;
    case CPV_VS_MVP_LAMBERT: p->vertex.host_fn = (const void*)&vertex_kernel<VsMvpLambert>; p->vs_outputs = VsMvpLambert::kOutputs; break;
    case CPV_VS_PASSTHROUGH_POINT:
        p->vertex.host_fn = (const void*)&vertex_kernel<VsPassthroughPoint>;
        p->vs_outputs = VsPassthroughPoint::kOutputs;
        p->psize = VsPassthroughPoint::kPointSize;
        break;
    default: return false;
    }
    switch (fs) {
    case CPV_FS_COLOR:
        for (int i = 0; i < kRasterVariants; i++) p->raster[i].host_fn = raster_table<FsColor>(i);
        p->fs_inputs = FsColor::kInputs;
        break;
    case CPV_FS_TEX_LAMBERT:
        for (int i = 0; i < kRasterVariants; i++) p->raster[i].host_fn = raster_table<FsTexLambert>(i);
        p->fs_inputs = FsTexLambert::kInputs;
        break;
    case CPV_FS_COLOR_DISCARD:
        for (int i = 0; i < kRasterVariants; i++) p->raster[i].host_fn = raster_table<FsColorDiscard>(i);
        p->fs_inputs = FsColorDiscard::kInputs;
        break;
    case CPV_FS_COLOR_FLAT:
        for (int i = 0; i < kRasterVariants; i++) p->raster[i].host_fn = raster_table<FsColorFlat>(i);
        p->fs_inputs = FsColorFlat::kInputs;
        break;
    case CPV_FS_COLOR_NOPERSPECTIVE:
        for (int i = 0; i < kRasterVariants; i++) p->raster[i].host_fn = raster_table<FsColorNoPerspective>(i);
        p->fs_inputs = FsColorNoPerspective::kInputs;
        break;
    default: return false;
    }
    return true;
}

int resolve_program(const cpv_pipeline_desc& pd, Program* p)
{
    if (pd.program) {
        const cpv_program* prog = pd.program;
        p->vertex = prog->vertex;
        for (int i = 0; i < kRasterVariants; i++) p->raster[i] = prog->raster[i];
        p->vs_outputs = prog->vs_outputs;
        p->psize = prog->psize;
        p->fs_inputs = prog->fs_inputs;
        return CPV_OK;
    }
    if (!builtin_program(pd.vs_kind, pd.fs_kind, p))
        return fail(CPV_ERROR_UNSUPPORTED, "unknown built-in program vs=%u fs=%u", pd.vs_kind, pd.fs_kind);
    return CPV_OK;
}

hipError_t launch(const KernelRef& k, dim3 grid, dim3 block, void* args, hipStream_t s)
{
    void* params[] = {args};
    if (k.host_fn) return hipLaunchKernel(k.host_fn, grid, block, params, 0, s);
    return hipModuleLaunchKernel(k.module_fn, grid.x, grid.y, grid.z, block.x, block.y, block.z, 0, s, params, nullptr);
}

// ------------------------------------------------------ descriptor build

int build_fragment_state(const cpv_pipeline_desc& p, const cpv_pass_desc& pass, FragmentState* fs)
{
    std::memset(fs, 0, sizeof(*fs));
    fs->has_color = pass.color_count > 0 && p.color_count > 0;
    fs->has_ds = pass.has_depth_stencil != 0;
    if (fs->has_color && !format_desc(pass.color[0].format, &fs->color))
        return fail(CPV_ERROR_UNSUPPORTED, "colour format %u", pass.color[0].format);
    if (fs->has_ds && !format_desc(pass.depth_stencil.format, &fs->ds))
        return fail(CPV_ERROR_UNSUPPORTED, "depth/stencil format %u", pass.depth_stencil.format);
    if (fs->has_color && is_depth_stencil(fs->color)) return fail(CPV_ERROR_INVALID, "depth format used as colour");
    if (fs->has_ds && !is_depth_stencil(fs->ds)) return fail(CPV_ERROR_INVALID, "colour format used as depth");
    // The reference aborts (TODO_ERROR) on these (PipelineCompiler.cpp:1674-1700,
    // CommandBuffer.Draw.cpp:1602-1605, 1675-1678, :1232-1236).
    if (p.logic_op_enable) return fail(CPV_ERROR_UNSUPPORTED, "logic op");
    if (p.rasterizer_discard) return fail(CPV_ERROR_UNSUPPORTED, "rasterizer discard");
    if (p.polygon_mode != CPV_POLYGON_FILL) return fail(CPV_ERROR_UNSUPPORTED, "polygon mode %u", p.polygon_mode);
    if (p.depth_clamp && p.depth_test) return fail(CPV_ERROR_UNSUPPORTED, "depth clamp");
    if (fs->has_color && p.blend[0].color_write_mask != CPV_COLOR_RGBA) return fail(CPV_ERROR_UNSUPPORTED, "colour write mask");
    if (fs->has_color && p.blend[0].blend_enable && (fs->color.base == kBaseUInt || fs->color.base == kBaseSInt))
        return fail(CPV_ERROR_UNSUPPORTED, "blend on an integer attachment");
    fs->depth_test = p.depth_test != 0;
    fs->depth_write = p.depth_write != 0;
    fs->depth_compare = (int32_t)p.depth_compare;
    fs->depth_bounds_test = p.depth_bounds_test != 0;
    fs->stencil_test = p.stencil_test != 0;
    fs->min_depth_bounds = p.min_depth_bounds;
    fs->max_depth_bounds = p.max_depth_bounds;
    if (p.color_count) {
        const cpv_blend_attachment& b = p.blend[0];
        fs->blend_enable = b.blend_enable != 0;
        fs->src_color = (int32_t)b.src_color_factor; fs->dst_color = (int32_t)b.dst_color_factor;
        fs->color_op = (int32_t)b.color_op;
        fs->src_alpha = (int32_t)b.src_alpha_factor; fs->dst_alpha = (int32_t)b.dst_alpha_factor;
        fs->alpha_op = (int32_t)b.alpha_op;
        if (b.blend_enable && (b.src_color_factor > CPV_BLEND_SRC_ALPHA_SATURATE || b.dst_color_factor > CPV_BLEND_SRC_ALPHA_SATURATE ||
                               b.src_alpha_factor > CPV_BLEND_SRC_ALPHA_SATURATE || b.dst_alpha_factor > CPV_BLEND_SRC_ALPHA_SATURATE ||
                               b.color_op > CPV_BLEND_OP_MAX || b.alpha_op > CPV_BLEND_OP_MAX))
            return fail(CPV_ERROR_UNSUPPORTED, "blend factor/op");
        fs->blend_kind = b.blend_enable && b.src_color_factor == CPV_BLEND_SRC_ALPHA && b.dst_color_factor == CPV_BLEND_ONE_MINUS_SRC_ALPHA &&
                                 b.color_op == CPV_BLEND_OP_ADD && b.src_alpha_factor == CPV_BLEND_SRC_ALPHA &&
                                 b.dst_alpha_factor == CPV_BLEND_ONE_MINUS_SRC_ALPHA && b.alpha_op == CPV_BLEND_OP_ADD
                             ? 1
                             : 0;
    }
    for (int i = 0; i < 4; i++) fs->blend_constants[i] = p.blend_constants[i];
    const cpv_stencil_state* st[2] = {&p.front, &p.back};
    int32_t* dst[2] = {fs->st_front, fs->st_back};
    for (int s = 0; s < 2; s++) {
        dst[s][0] = (int32_t)st[s]->fail_op; dst[s][1] = (int32_t)st[s]->pass_op;
        dst[s][2] = (int32_t)st[s]->depth_fail_op; dst[s][3] = (int32_t)st[s]->compare_op;
        dst[s][4] = (int32_t)st[s]->compare_mask; dst[s][5] = (int32_t)st[s]->write_mask;
        dst[s][6] = (int32_t)st[s]->reference;
    }
    fs->origin_upper_left = p.origin_upper_left != 0;
    return CPV_OK;
}

int build_texture(const cpv_texture_desc& t, TextureDesc* o)
{
    std::memset(o, 0, sizeof(*o));
    if (!t.address) return CPV_OK;
    if (!format_desc(t.format, &o->fmt) || is_depth_stencil(o->fmt)) return fail(CPV_ERROR_UNSUPPORTED, "texture format %u", t.format);
    if (t.levels == 0 || t.levels > 16 || t.base_level + t.level_count > t.levels || t.level_count == 0)
        return fail(CPV_ERROR_INVALID, "texture levels");
    o->address = t.address;
    o->level_count = t.level_count;
    // GetNormalImageSize (CPVulkanBase/Formats.cpp:455-482)
    uint64_t off = 0;
    uint32_t w = t.width, h = t.height;
    for (uint32_t i = 0; i < t.levels; i++) {
        if (i >= t.base_level && i < t.base_level + t.level_count) {
            const uint32_t j = i - t.base_level;
            o->level_offset[j] = off;
            o->level_w[j] = w;
            o->level_h[j] = h;
        }
        off += (uint64_t)o->fmt.total * w * h;
        w = std::max(w / 2, 1u);
        h = std::max(h / 2, 1u);
    }
    const cpv_sampler_desc& s = t.sampler;
    if (s.unnormalized) return fail(CPV_ERROR_UNSUPPORTED, "unnormalised coordinates");
    o->sampler.mag_filter = (int32_t)s.mag_filter;
    o->sampler.min_filter = (int32_t)s.min_filter;
    o->sampler.mipmap_mode = (int32_t)s.mipmap_mode;
    o->sampler.address_u = (int32_t)s.address_u;
    o->sampler.address_v = (int32_t)s.address_v;
    o->sampler.mip_lod_bias = s.mip_lod_bias;
    o->sampler.min_lod = s.min_lod;
    o->sampler.max_lod = s.max_lod;
    o->sampler.border_color = (int32_t)s.border_color;
    return CPV_OK;
}

int build_env(const cpv_pipeline_desc& p, const cpv_draw_desc& d, DrawEnv* e)
{
    std::memset(e, 0, sizeof(*e));
    for (uint32_t i = 0; i < p.attribute_count && i < CPV_MAX_VERTEX_ATTRIBUTES; i++) {
        const cpv_vertex_attribute_desc& a = p.attributes[i];
        if (a.location >= 16) return fail(CPV_ERROR_INVALID, "attribute location %u", a.location);
        AttribDesc& o = e->attr[a.location];
        if (!format_desc(a.format, &o.fmt)) return fail(CPV_ERROR_UNSUPPORTED, "attribute format %u", a.format);
        o.binding = a.binding;
        o.offset = a.offset;
        o.valid = 1;
    }
    for (uint32_t i = 0; i < p.binding_count && i < CPV_MAX_VERTEX_BINDINGS; i++) {
        const cpv_vertex_binding_desc& b = p.bindings[i];
        if (b.binding >= 16) return fail(CPV_ERROR_INVALID, "binding %u", b.binding);
        BindingDesc& o = e->bind[b.binding];
        o.address = d.vertex_address[b.binding];
        o.bytes = d.vertex_bytes[b.binding];
        o.stride = b.stride;
        o.input_rate = b.input_rate;
    }
    for (int i = 0; i < 4; i++) e->uniform[i] = d.uniform_address[i];
    for (int i = 0; i < 4; i++) {
        int rc = build_texture(d.textures[i], &e->tex[i]);
        if (rc) return rc;
    }
    return CPV_OK;
}

int build_target(const cpv_pass_desc& pass, bool first_segment, int tile, uint32_t tile_base, TargetParams* t)
{
    std::memset(t, 0, sizeof(*t));
    if (pass.width == 0 || pass.height == 0 || pass.width > 16384 || pass.height > 16384)
        return fail(CPV_ERROR_INVALID, "pass extent %ux%u", pass.width, pass.height);
    t->width = pass.width;
    t->height = pass.height;
    t->tiles_x = (pass.width + tile - 1) / tile;
    t->tiles_y = (pass.height + tile - 1) / tile;
    t->tile_base = tile_base;
    t->has_color = pass.color_count > 0 && pass.color[0].address != 0;
    t->has_ds = pass.has_depth_stencil && pass.depth_stencil.address != 0;
    FormatDesc f;
    if (t->has_color) {
        if (!format_desc(pass.color[0].format, &f)) return fail(CPV_ERROR_UNSUPPORTED, "colour format %u", pass.color[0].format);
        t->color_fmt = f;
        t->color_addr = pass.color[0].address;
        t->color_bytes = f.total;
        // DONT_CARE keeps memory as is in the reference, which is LOAD here.
        t->color_load = (first_segment && pass.color[0].load_op == CPV_LOAD_OP_CLEAR) ? 1 : 0;
        for (int i = 0; i < 4; i++) t->clear_color[i] = pass.color[0].clear.u32[i];
    }
    if (t->has_ds) {
        if (!format_desc(pass.depth_stencil.format, &f)) return fail(CPV_ERROR_UNSUPPORTED, "depth format %u", pass.depth_stencil.format);
        t->ds_fmt = f;
        t->ds_addr = pass.depth_stencil.address;
        t->ds_bytes = f.total;
        t->ds_load = (first_segment && pass.depth_stencil.load_op == CPV_LOAD_OP_CLEAR) ? 1 : 0;
        std::memcpy(&t->clear_ds[0], &pass.depth_stencil.clear.ds.depth, 4);
        t->clear_ds[1] = pass.depth_stencil.clear.ds.stencil;
    }
    return CPV_OK;
}

// CalculatePrimitives (CommandBuffer.Draw.cpp:567-673).
uint32_t prim_count_of(uint32_t topology, uint32_t n)
{
    if (topology == CPV_TOPOLOGY_POINT_LIST) return n;
    if (topology == CPV_TOPOLOGY_LINE_LIST) return n / 2;
    if (topology == CPV_TOPOLOGY_LINE_STRIP) return n > 1 ? n - 1 : 0;
    if (topology == CPV_TOPOLOGY_TRIANGLE_LIST) return n / 3;
    if (topology == CPV_TOPOLOGY_TRIANGLE_STRIP || topology == CPV_TOPOLOGY_TRIANGLE_FAN) return n > 2 ? n - 2 : 0;
    return 0;
}

}  // namespace cpvrt

extern "C" int cpv_abi_sizes(uint32_t* out, uint32_t n)
{
    const uint32_t sizes[] = {(uint32_t)sizeof(cpv_pipeline_desc), (uint32_t)sizeof(cpv_draw_desc), (uint32_t)sizeof(cpv_pass_desc),
                              (uint32_t)sizeof(cpv_attachment_desc), (uint32_t)sizeof(cpv_texture_desc), (uint32_t)sizeof(cpv_sampler_desc),
                              (uint32_t)sizeof(cpv_job_desc), (uint32_t)sizeof(cpv_config), (uint32_t)sizeof(cpv_tile_trace),
                              (uint32_t)sizeof(cpv_autotune_entry), (uint32_t)sizeof(cpv_copy_region), (uint32_t)sizeof(cpv_image_level),
                              (uint32_t)sizeof(cpv_rect), (uint32_t)sizeof(cpv_blit_region)};
    const uint32_t m = (uint32_t)(sizeof(sizes) / sizeof(sizes[0]));
    for (uint32_t i = 0; i < n && i < m; i++) out[i] = sizes[i];
    return (int)m;
}

// ------------------------------------------------------- SPIR-V programs

int cpv_spirv_compile(cpv_context* c, const uint32_t* vs, size_t vs_n, const char* vs_entry, const uint32_t* fs, size_t fs_n,
                      const char* fs_entry, cpv_program** out, std::string* err);

extern "C" int cpv_program_create_spirv(cpv_context* c, const uint32_t* vs_words, size_t vs_word_count, const char* vs_entry,
                                        const uint32_t* fs_words, size_t fs_word_count, const char* fs_entry, cpv_program** out)
{
    if (!c || !vs_words || !fs_words || !out) return fail(CPV_ERROR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    std::string err;
    int rc = cpv_spirv_compile(c, vs_words, vs_word_count, vs_entry, fs_words, fs_word_count, fs_entry, out, &err);
    if (rc) return fail(rc, "%s", err.c_str());
    return CPV_OK;
}

extern "C" void cpv_program_destroy(cpv_context* c, cpv_program* p)
{
    if (!p) return;
    if (c) (void)hipSetDevice(c->device);
    if (p->module) (void)hipModuleUnload(p->module);
    delete p;
}

extern "C" const char* cpv_program_source(cpv_program* p) { return p ? p->source.c_str() : ""; }
