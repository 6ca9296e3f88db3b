// Host-side declarations shared by the translation units of libcpv.so
// (runtime, launch sequence, tile-cost predictor).  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cpv.h"
#include "cpv_program.h"
#include "cpv_types.h"
#include "raster_device.h"

namespace cpvrt {

int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                                            \
    do {                                                                                                         \
        hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess) return ::cpvrt::fail(CPV_ERROR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

// Grow-only device buffer.  With `frozen` set (while a graph is captured)
// growing is an error instead of an allocation.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes, bool frozen = false);
    template <class T> T* as() const { return (T*)p; }
    void release();
};

// Control words at the start of Buffers::stats.
enum : uint32_t {
    kStatFragments = 0,   // u64 at words 0-1: covered fragments
    kStatPairs = 9,       // (tile, primitive) pairs kept, min(total, capacity)
    // set when the pair total exceeded the capacity; on a 128-byte line of
    // its own, since every raster workgroup reads it at its start while the
    // waves add to the fragment counter at word 0 (sharing that line made
    // the raster kernel 30% slower)
    kStatOverflow = 32,
    kStatTiles = 11,      // tile count of the launch (length of the order sort)
    kStatPrio = 12,       // predicted log2 cost from which a tile is heavy and dispatched first (float bits)
    kStatPredMax = 13,    // the largest finite prediction, as an ordered integer
    kStatPairTotal = 14,  // u64 at words 14-15: (tile, primitive) pairs the launch's primitives need
    kStatWords = 64,      // words read back after a launch
    // Work counters of the persistent raster kernel, one per queue, each on
    // its own 128-byte line, after the read-back words.
    kStatQueues = kStatWords,
    kStatAlloc = kStatWords + cpv::kRasterQueues * cpv::kRasterQueueStride,
};

// Working buffers of the launch sequence.  A context owns one set for its
// immediate render calls; a recorded batch owns its own so that replaying
// it never reallocates.
struct Buffers {
    DevBuf vout, recs, counts, offsets, keys[2], vals[2], tile_start, hist, hist_base, sums, big;
    DevBuf arena, stats, traces, features, pred, okeys[2], ovals[2], line_ext;
    void release_all();
};

struct Program {
    KernelRef vertex;
    KernelRef raster[cpv::kRasterVariants];
    int vs_outputs = 0;
    int psize = -1;       // vertex output component holding gl_PointSize, -1: none
    int fs_inputs = 0;
};

struct MlpDevice;   // the tile-cost predictor's device state (cpv_mlp.hip)

// The last immediate launch of a context, kept until its pair count has been
// checked (cpv_launch.hip): the launch ran with a pair capacity taken from
// earlier launches, and the next call on the context redoes it with a larger
// one if its pairs did not fit (its raster kernel then drew nothing).
struct PendingLaunch;

// A byte range of device memory [lo, hi).
struct Span {
    uint64_t lo = 0, hi = 0;
    bool meets(const Span& o) const { return lo < o.hi && o.lo < hi; }
};
void destroy_pending(PendingLaunch* p);
int resolve_pending(cpv_context* c, bool may_allocate = true);
// Records a recorded batch's submit again after it outgrew its pair arrays
// (cpv_launch.hip), and re-records the fences signalled after the redrawn
// work.
int batch_redo(cpv_context* c, cpv_batch* B, uint32_t set, bool may_allocate);
void rerecord_fences(cpv_context* c);

// Takes the device address of this GPU's scratch array (the working memory
// of cpv_render_pass_noalloc) into c->scratch; once per device.
int scratch_setup(cpv_context* c);
size_t scratch_bytes();

}  // namespace cpvrt

struct cpv_mlp;

struct cpv_context {
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    hipStream_t front_stream = nullptr;   // front halves (vertex, setup, binning) of launches and batch submits
    uint64_t enqueue_seq = 0;             // calls that enqueued work on the context (cpvrt::note_enqueue)
    hipEvent_t ev_begin = nullptr, ev_raster0 = nullptr, ev_raster1 = nullptr, ev_end = nullptr;
    cpv_config cfg{};
    cpvrt::Buffers buf;                   // the render calls' control words (fragment counter)
    // Working buffers of immediate launches, used alternately: a launch's
    // front half on front_stream waits only for the raster kernel that last
    // used its set, so it runs under the previous launch's raster kernel.
    cpvrt::Buffers lbuf[2];
    hipEvent_t set_front_done[2] = {}, set_raster_done[2] = {};
    bool set_used[2] = {false, false};
    uint32_t next_set = 0, last_set = 0;
    uint64_t chain_seq = ~0ull;           // enqueue_seq at the last launch: equal means nothing came in between
    hipEvent_t ev_pre = nullptr;          // joins front_stream to the context stream's earlier work
    bool call_first_launch = false;       // the render call has not launched yet (its begin event is pending)
    std::vector<cpvrt::Span> last_written;   // attachment bytes the last immediate launch draws into
    // Pinned staging of host-to-device uploads (launch descriptors,
    // cpv_update_buffer data): a ring of slots, each written again only once
    // the copy issued from it has completed, so an upload does not wait for
    // the work enqueued before it.
    struct Staged {
        void* p = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;   // recorded after the copy issued from the slot
        bool used = false;
    };
    Staged staged[4];
    uint32_t staged_next = 0;
    uint32_t* readback = nullptr;     // pinned scalar readbacks
    // what the last render call left behind (for traces / counts)
    uint32_t last_tiles = 0;
    uint32_t last_waves = 0;
    uint32_t last_tile_size = 0;
    std::vector<uint32_t> last_target_base;   // first tile of every target of the last launch
    bool last_traces = false;
    bool timing_valid = false;
    bool last_count_exact = false;
    cpv_mlp* predictor = nullptr;     // tile-cost model of CPV_SCHED_PREDICTED
    std::map<const void*, int> occupancy;   // resident workgroups per CU, by raster kernel
    cpvrt::PendingLaunch* pending = nullptr;   // the last immediate launch, until its pair count is checked
    hipEvent_t ev_check = nullptr;             // recorded after the pending launch's pair-count readback
    uint64_t pair_cap = 0;                     // pair capacity of immediate launches (grows as launches need)
    uint8_t* scratch = nullptr;                // device address of this GPU's scratch array (cpvrt::scratch_setup)
    uint32_t* last_stats = nullptr;            // control words of the last render call (fragment count)
    uint64_t launches = 0;                     // raster launches enqueued: render calls and batch submits
    cpv_batch* pending_batch = nullptr;        // the last batch submit, until its overflow word is checked
    uint32_t pending_set = 0;                  // the set of working buffers that submit used
    std::vector<cpv_fence*> fences_after_pending;   // fences signalled since the pending launch or submit
    std::mutex lock;
};

namespace cpvrt {

bool format_desc(uint32_t f, cpv::FormatDesc* o);

// Host address of raster kernel variant i of built-in fragment shader FS
// (instantiated by raster_instances.hip, one translation unit per shader).
template <class FS>
const void* raster_table(int i);
int tile_config_index(uint32_t t, uint32_t w);
bool valid_tile_config(uint32_t t, uint32_t w);
bool builtin_program(uint32_t vs, uint32_t fs, Program* p);
int resolve_program(const cpv_pipeline_desc& pd, Program* p);
hipError_t launch(const KernelRef& k, dim3 grid, dim3 block, void* args, hipStream_t s);
int build_fragment_state(const cpv_pipeline_desc& p, const cpv_pass_desc& pass, cpv::FragmentState* fs);
int build_env(const cpv_pipeline_desc& p, const cpv_draw_desc& d, cpv::DrawEnv* e);
int build_target(const cpv_pass_desc& pass, bool first_segment, int tile, uint32_t tile_base, cpv::TargetParams* t);
uint32_t prim_count_of(uint32_t topology, uint32_t n);
void fill_u32(hipStream_t s, void* p, uint32_t v, uint64_t n);
// Copies bytes from host memory to device address dst on the context's
// stream through the next pinned staging slot; src is free on return.
// Without may_grow a slot too small for the bytes is an error, not an
// allocation.
int stage_upload(cpv_context* c, void* dst, const void* src, size_t bytes, bool may_grow = true, hipStream_t s = nullptr);
// Counts a call that enqueued work on the context's stream: a recorded
// batch overlaps a submit with its previous one only when nothing else was
// enqueued in between.
inline void note_enqueue(cpv_context* c) { c->enqueue_seq++; }
// Waits for all work of the context, its batches' front halves included.
inline hipError_t sync_context(cpv_context* c)
{
    const hipError_t e = c->front_stream ? hipStreamSynchronize(c->front_stream) : hipSuccess;
    return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}

// Tile-cost predictor (cpv_mlp.hip): enqueue inference of `tiles` feature
// rows (16 floats each) into per-tile predicted log2 costs.
int mlp_enqueue_predict(cpv_mlp* m, const float* features, uint32_t tiles, float* out, hipStream_t s);
constexpr size_t kStagingSlotBytes = 256u << 10;   // initial size of each pinned staging slot
constexpr int kFeatures = 12;       // features per tile the predictor reads
constexpr int kFeatureStride = 16;  // floats per feature row (padded to the MFMA depth)

}  // namespace cpvrt
