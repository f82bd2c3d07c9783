// context.h -- host-side state of a libgudni_hip context, shared by the
// translation units that implement the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "gudni_hip.h"
#include "internal.h"
#include "device.h"

namespace gudni {

// A device allocation that only grows.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n)
    {
        if (n <= bytes && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        const size_t want = n ? n : 16;
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

} // namespace gudni


struct gudni_ctx {
    gudni_spec spec;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
    int cus = 0;
    int fast_blocks = 0, slow_blocks = 0;
    int blocks_per_cu = 0;        // fast-pass blocks per CU (0: as many as fit, at most 8)
    int32_t out_format = GUDNI_FORMAT_BGRA8;   // pixel word byte order
    // common buffers
    gudni::DevBuf geo, subs, pict, uses, random, grads;
    int32_t n_substances = 0, n_uses = 0, n_grads = 0;
    // host copies of the common buffers, used to bounds-check job records
    std::vector<uint8_t> geo_host;
    std::vector<float4> subs_host;
    // per launch buffers
    gudni::DevBuf tiles, tile_job, shapes, xrange, work, out, out_f32;
    gudni::DevBuf fast_thr, fast_hdr, fast_sidx, slow_thr, slow_hdr, slow_sidx;
    gudni::DevBuf control;   // [0] fast counter [1] slow counter [2] status [3..] retry
    gudni::DevBuf scene_table;           // single-frame scene descriptor
    gudni::dev::SceneDesc single_scene;  // host copy of what scene_table holds
    gudni::DevBuf cycles;    // stage cycle counters of the fast pass (timing only)
    gudni::DevBuf dump;      // dump staging
    // pinned host staging of the frame box a Level-1 job reads and writes
    void *l1_stage = nullptr;
    size_t l1_stage_bytes = 0;
    int32_t retry_capacity = 0;
    // bumped whenever the launch scratch (rings, control words) or the fast
    // grid may have changed: a captured graph holds those pointers and sizes
    uint64_t scratch_gen = 0;
    // status word of the last launch (the context's control words, or the
    // library's static scratch after gudni_frame_render_static)
    uint32_t *status_word = nullptr;
    // uploaded frame (Level 2)
    int32_t frame_w = 0, frame_h = 0, frame_n_work = 0, frame_n_tiles = 0, frame_index = 0;
    int32_t frame_sort_bits = 0;  // sort radix of the spec the uploaded jobs were built with
    float background[4] = {1, 1, 1, 1};
    bool frame_ready = false;
    // Level 2 scene: shapes in traversal order, binned on the device
    gudni::DevBuf scene_shapes, scene_bbox, scene_bounds, bin_ws;
    gudni::DevBuf bins_ws, bins_jobs;   // many-scene binning: workspaces, job records
    int32_t scene_n_shapes = 0, scene_w = 0, scene_h = 0;
    float scene_bg[4] = {1, 1, 1, 1};
    int32_t bin_cap_pairs = 0, bin_cap_nodes = 0, scene_n_refs = 0;
    bool scene_ready = false;
    hipEvent_t bin_ev[2] = {nullptr, nullptr};
    float last_bin_ms = 0;
    // device serialisation (gudni_scene_upload_outlines)
    gudni::DevBuf ser_tags, ser_soff, ser_ooff, ser_pairs, ser_work;
    gudni::DevBuf ser_many_ws, ser_many_args;   // many-scene serialisation
    size_t scene_geo_bytes = 0;
    hipEvent_t ser_ev[2] = {nullptr, nullptr};
    float last_ser_ms = 0;
    // timing
    bool timing = false, timed_render = false, timed_adjoint = false;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    float last_ms[2] = {0, 0};
    unsigned long long last_cycles[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

#define CHECK_HIP(ctx, expr)                                                        \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            (ctx)->error = std::string(#expr) + ": " + hipGetErrorString(e_);      \
            return GUDNI_ERR_HIP;                                                   \
        }                                                                           \
    } while (0)

namespace gudni {

inline int set_error(gudni_ctx *ctx, int code, const std::string &msg)
{
    ctx->error = msg;
    return code;
}

template <typename T>
inline hipError_t upload(gudni_ctx *ctx, DevBuf &b, const T *src, size_t bytes)
{
    hipError_t e = b.ensure(bytes);
    if (e != hipSuccess) return e;
    if (bytes && src) return hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    return hipSuccess;
}

// Host copies of the common buffers, for record validation.
struct CommonView {
    const uint8_t *geo;
    size_t geo_bytes;
    const float4 *subs;
    int32_t n_subs, n_uses, n_grads;
};

// Checks gradient records (kind 0 or 1, 1..4 stops).
int validate_gradients(gudni_ctx *ctx, const void *grads, int32_t n);

CommonView ctx_view(gudni_ctx *ctx);
// Active (tile, quarter) work units of a tile list (isActiveThread).
void build_work(const gudni_spec &spec, const TileRec *tiles, int n_tiles, int width, int height,
                std::vector<uint32_t> &work);
// The work list of a whole-batch launch: the long units (estimated cost
// rows per lane x ((tile shapes + 1) / tile width)^2 within a factor 8 of the
// longest) sorted longest first and dealt tile by tile, round robin, to the
// starts of the `parts` contiguous ranges the claimers begin with, so none is
// claimed last.  With tile_scene (each tile's scene) the other units follow
// scene by scene, the cheapest scene (sum of its units' costs) first, each
// scene's units together and dealt round robin to the ranges, so scenes
// finish in order of cost on every XCD; without it they keep their original
// order, range after range.  Either way a scene's tiles stay on one XCD.
void order_longest_first(int threads_per_tile, const TileRec *tiles, const int32_t *tile_scene,
                         const uint32_t *work, int n, int parts, uint32_t *out);
// Scratch rings of the raster passes, sized on first use.
int ensure_scratch(gudni_ctx *ctx);
// Pass counters, status word and retry list for n_work units.
int ensure_control(gudni_ctx *ctx, int n_work);
// A Frame with the context's spec and buffers; the caller sets the rest.
dev::Frame base_frame(gudni_ctx *ctx);
// Waits for the stream and turns the status word into an error code.
int read_status(gudni_ctx *ctx);
int validate_uses(gudni_ctx *ctx, const void *pict_uses, int32_t n_pict_uses, size_t pict_bytes);
// Stage (a) on the device for one scene whose shape records, boxes and
// strand extents are in device memory: tiles, reversed job index per tile,
// shape references, their extents and the work list into the caller's
// buffers (grown as needed), with the context's spec.  One
// synchronisation (the counts); the capacities retry when exceeded.
struct BinIn {
    const ShapeRec *shapes;
    const float4 *bbox, *bounds;
    int32_t n_shapes, width, height;
    int32_t max_tile_size = 0;     // 0: the spec's
};
struct BinOut {
    DevBuf *tiles, *tile_job, *refs, *ref_bounds, *work;
};
struct BinCounts {
    int32_t n_tiles = 0, n_refs = 0, n_work = 0;
};
int bin_scene(gudni_ctx *ctx, const BinIn &in, const BinOut &out, BinCounts &cnt);
// The same for n scenes at once, one block per scene (for scenes of a few
// thousand shapes, where the per-level launches of bin_scene cost more than
// the work): one launch and one synchronisation while the capacities hold.
int bin_scenes(gudni_ctx *ctx, const BinIn *in, const BinOut *out, BinCounts *cnt, int n);
// Bounds-checks shape records (and tiles) against the common buffers and
// fills each shape's strand extent (xmin, ymin, xmax, ymax).
int validate_job(gudni_ctx *ctx, const CommonView &cv, const ShapeRec *shapes, int32_t n_shapes,
                 const TileRec *tiles, int32_t n_tiles, int32_t w, int32_t h,
                 std::vector<float> &xrange);

} // namespace gudni
