// raster.hip -- gfx950 kernels and C ABI of libgudni_hip.
//
// One lane owns one column segment of a tile (the reference's work-item,
// Kernels.cl:1692-1722) and runs threshold generation, the per-column sort
// and the scanline coverage/compositing sweep back to back, so records never
// leave the wave's scratch between stages.  Work is claimed by whole
// wavefronts: a work unit (tile << 2 | q) is a block of 64 of the tile's
// 256 column segments, 8 columns x 8 segments where the tile allows
// (unit_column), and units without an active column are never queued (an
// 8x8 tile is a single unit).  The grid is persistent; every wave loops on an atomic
// work counter until the list is exhausted, so every wave reaches the exit.
//
// Each lane's threshold queue is a ring of its own in the wave slot's
// scratch.  The fast pass gives each lane a ring of kFastCap records; a lane
// whose live queue outgrows it abandons its unit, which is re-run by a second
// persistent pass whose rings hold maxThresholds records.  The result is
// identical to one pass at full capacity, with ~1/32 of the scratch.
//
#include <hip/hip_runtime.h>

#include <mutex>
#include <system_error>
#include <thread>
#include <unordered_map>

#include <cstring>
#include <string>
#include <vector>

#include "bin.h"
#include "context.h"
#include "device.h"
#include "serialize.h"

namespace gudni {
namespace dev {

// One persistent pass.  A wave claims a work unit, its lanes run their
// columns to completion, and the wave loops until the list is empty.
// The column walks are bound by the latency of their ring and strand reads,
// so residency pays: five waves per SIMD (at most 96 VGPRs, a few spills)
// render the 4096^2 frame ~25% faster than four and ~30% faster than three.
// Five blocks' tile tables (4 x 7.2 KB each) fit the CU's 160 KB of LDS.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) raster_pass(Frame f, int from_retry)
{
    __shared__ TileTable tables[4];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int slot = blockIdx.x * (blockDim.x / kWave) + wave;
    const int cap = 1 << f.cap_log2;
    float4 *q_thr = f.q_thr + ((size_t)slot * kWave + lane) * cap;
    uint32_t *q_hdr = f.q_hdr + ((size_t)slot * kWave + lane) * cap;
    int32_t *sidx = f.shape_idx + (size_t)slot * kSidxWords * kWave + lane;
    TileTable *tt = &tables[wave];
    const int n_work = from_retry ? (int)min((uint32_t)f.retry_capacity, f.retry[0]) : f.n_work;
    Claimer claim;
    claim.init(f.counter, f.n_parts, n_work, (int)blockIdx.x);
    for (;;) {
        const int w = claim.next(lane);
        if (w < 0) break;
        const uint32_t unit = from_retry ? f.retry[1 + w] : f.work[w];
        const int count = f.tiles[unit >> 2].shape_count;   // wave-uniform
        const int scene = f.tile_scene
                              ? (int)__builtin_amdgcn_readfirstlane((uint32_t)f.tile_scene[unit >> 2])
                              : 0;
        // wave-uniform, read through the scalar cache where a field is used
        // (a register copy of the whole descriptor spilled 45 more scalar
        // registers; the 4096^2 frame: 43.1 -> 42.7 ms)
        const SceneDesc &s = *(const SceneDesc *)ro(f.scenes + scene);
        Queue q;
        q.thr = q_thr; q.hdr = q_hdr;
        q.head = 0; q.len = 0; q.mask = cap - 1; q.maxt = f.max_thresholds;
        q.spilled = false; q.overflow = false;
        uint32_t status = 0;
        if (count <= f.max_shapes) run_unit<true>(f, s, lane, unit, q, sidx, tt, status);
        else run_unit<false>(f, s, lane, unit, q, sidx, tt, status);
        if (status) atomicOr(f.status, status);
        if (f.scene_done && lane == 0)
            atomicMax(f.scene_done + scene, (unsigned long long)wall_clock64());
        // a lane whose ring was too small hands the whole unit to the retry pass
        const bool spilled = __any(q.spilled);
        if (spilled && lane == 0 && !from_retry) {
            const uint32_t k = atomicAdd(f.retry, 1u);
            if ((int)k < f.retry_capacity) f.retry[1 + k] = unit;
            else atomicOr(f.status, kStatusOverflow);
        }
    }
}

__global__ void reset_counters(uint32_t *counters, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counters[i] = 0u;
}

// Writes a scene descriptor passed by value (no host-to-device copy).
__global__ void store_scene(SceneDesc *dst, SceneDesc sd)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = sd;
}

// Start of a batch: clears the per-scene completion times and stamps the
// start time in done[0].
__global__ void batch_begin(unsigned long long *done, int n_scenes)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) done[0] = wall_clock64();
    else if (i <= n_scenes) done[i] = 0ull;
}

} // namespace dev
} // namespace gudni

// ===========================================================================
// The static scratch of gudni_frame_render_static
// ===========================================================================
// Callers that cannot allocate device memory (code under graph capture,
// serving loops) render an uploaded frame with rings, shape scratch, control
// words and scene descriptor carved from this one array, which every GPU the
// library runs on holds until the process ends.  Layout (bytes):
//   fast pass, 32 blocks = 128 wave slots: threshold rings of 128 records
//     (16 MiB of float4, 4 MiB of headers) and shape scratch (8 MiB)
//   retry pass, one wave slot: a ring of 4096 records (4 MiB + 1 MiB) and
//     shape scratch (64 KiB)
//   control words and a retry list of up to 2^17 work units (512 KiB)
//   one scene descriptor
namespace {
constexpr int kStaticFastBlocks = 32;
constexpr size_t kStaticFastSlots = (size_t)kStaticFastBlocks * 4;
constexpr int kStaticSlowLog2 = 12;
constexpr int kStaticMaxWork = 1 << 17;
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kStaticFastRecs = (size_t)1 << gudni::dev::kFastCapLog2;
constexpr size_t kStaticSlowRecs = (size_t)1 << kStaticSlowLog2;
constexpr size_t kOffFastThr = 0;
constexpr size_t kOffFastHdr = kOffFastThr + kStaticFastSlots * kStaticFastRecs * 64 * 16;
constexpr size_t kOffFastSidx = kOffFastHdr + kStaticFastSlots * kStaticFastRecs * 64 * 4;
constexpr size_t kOffSlowThr = kOffFastSidx + kStaticFastSlots * gudni::dev::kSidxWords * 64 * 4;
constexpr size_t kOffSlowHdr = kOffSlowThr + kStaticSlowRecs * 64 * 16;
constexpr size_t kOffSlowSidx = kOffSlowHdr + kStaticSlowRecs * 64 * 4;
constexpr size_t kOffCtl = kOffSlowSidx + (size_t)gudni::dev::kSidxWords * 64 * 4;
constexpr size_t kOffScene = align256(kOffCtl + 4 * ((size_t)gudni::dev::kCtlWords + kStaticMaxWork));
constexpr size_t kStaticBytes = align256(kOffScene + sizeof(gudni::dev::SceneDesc));
static_assert(kStaticBytes <= ((size_t)34 << 20), "the static scratch is stated as at most 34 MiB");
constexpr int kStaticMaxDevices = 64;
}  // namespace

static __device__ __align__(256) uint8_t scratch[kStaticBytes];

namespace {
// Device address of `scratch` on each device (each GPU holds its own copy),
// looked up once per device by gudni_create with that device current.
std::mutex g_scratch_mu;
void *g_scratch_ptr[kStaticMaxDevices];
bool g_scratch_looked_up[kStaticMaxDevices];
hipError_t g_scratch_err[kStaticMaxDevices];

void lookup_static_scratch(int device)
{
    if (device < 0 || device >= kStaticMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_looked_up[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    g_scratch_looked_up[device] = true;
}

uint8_t *static_scratch(int device, hipError_t &err)
{
    err = hipErrorInvalidDevice;
    if (device < 0 || device >= kStaticMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_looked_up[device]) return nullptr;
    err = g_scratch_err[device];
    return (uint8_t *)g_scratch_ptr[device];
}
}  // namespace

// ===========================================================================
// Host side of the device API
// ===========================================================================
using namespace gudni;

namespace gudni {

// Which work units of a tile have an active column (isActiveThread,
// Kernels.cl:1720-1722), evaluated for every column on the host.
void build_work(const gudni_spec &spec, const TileRec *tiles, int n_tiles, int width,
                       int height, std::vector<uint32_t> &work)
{
    const int cd = ceil_log2(spec.threads_per_tile);
    const int quarters = spec.threads_per_tile / 64;
    for (int t = 0; t < n_tiles; ++t) {
        const TileRec &r = tiles[t];
        const int hd = r.hdepth, vd = r.vdepth;
        const int dd = std::max(0, vd - (cd - hd));
        for (int qd = 0; qd < quarters; ++qd) {
            bool any = false;
            for (int l = 0; l < 64 && !any; ++l) {
                const int col = unit_column(hd, cd, qd, l);
                const int ix = ((1 << hd) - 1) & col, iy = (col >> hd) << dd;
                any = iy < (r.box[3] - r.box[1]) && ix + r.box[0] < width && iy + r.box[1] < height;
            }
            if (any) work.push_back(((uint32_t)t << 2) | (uint32_t)qd);
        }
    }
}

void order_longest_first(int threads_per_tile, const TileRec *tiles, const int32_t *tile_scene,
                         const uint32_t *work, int n, int parts, uint32_t *out)
{
    const int cd = ceil_log2(threads_per_tile);
    // a lane sweeps `rows` pixels whose cost grows with the tile's shape
    // density per column: rows x ((shapes + 1) / width)^2
    std::vector<double> cost((size_t)n);
    double top = 0.0;
    for (int i = 0; i < n; ++i) {
        const TileRec &r = tiles[work[i] >> 2];
        const int rows = std::max(1, (r.box[3] - r.box[1]) >> std::max(0, cd - (int)r.hdepth));
        const double d = (double)(std::max(0, r.shape_count) + 1) / (double)(1 << std::max(0, (int)r.hdepth));
        cost[i] = rows * d * d;
        top = std::max(top, cost[i]);
    }
    // the long units: within kSpread of the longest
    constexpr double kSpread = 8.0;
    std::vector<int> heavy;
    for (int i = 0; i < n; ++i)
        if (cost[i] * kSpread >= top) heavy.push_back(i);
    std::stable_sort(heavy.begin(), heavy.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<long long> begin((size_t)parts), size((size_t)parts), fill((size_t)parts, 0);
    for (int p = 0; p < parts; ++p) {
        begin[p] = (long long)n * p / parts;
        size[p] = (long long)n * (p + 1) / parts - begin[p];
    }
    // long units first, dealt tile by tile round robin to the range starts
    std::vector<char> placed((size_t)n, 0);
    int p = 0;
    for (size_t i = 0; i < heavy.size();) {
        size_t j = i + 1;
        while (j < heavy.size() && (work[heavy[j]] >> 2) == (work[heavy[i]] >> 2)) ++j;
        for (size_t k = i; k < j; ++k) {
            for (int tries = 0; fill[p] == size[p] && tries < parts; ++tries) p = (p + 1) % parts;
            out[begin[p] + fill[p]++] = work[heavy[k]];
            placed[heavy[k]] = 1;
        }
        p = (p + 1) % parts;
        i = j;
    }
    if (!tile_scene) {
        // then every other unit in its original order, range after range
        p = 0;
        for (int i = 0; i < n; ++i) {
            if (placed[i]) continue;
            while (fill[p] == size[p]) ++p;
            out[begin[p] + fill[p]++] = work[i];
        }
        return;
    }
    // then the scenes' other units, cheapest scene first, a scene at a time
    // dealt round robin to the ranges (its units together and in their
    // order, spilling into the next range with room when one is full): each
    // XCD finishes its small scenes early instead of every scene finishing
    // near the end of the batch
    std::vector<int> scenes;
    std::unordered_map<int, std::vector<int>> units_of;
    std::unordered_map<int, double> scene_cost;
    for (int i = 0; i < n; ++i) {
        if (placed[i]) continue;
        const int sc = tile_scene[work[i] >> 2];
        auto it = units_of.find(sc);
        if (it == units_of.end()) {
            scenes.push_back(sc);
            it = units_of.emplace(sc, std::vector<int>()).first;
        }
        it->second.push_back(i);
        scene_cost[sc] += cost[i];
    }
    std::stable_sort(scenes.begin(), scenes.end(), [&](int a, int b) { return scene_cost[a] < scene_cost[b]; });
    p = 0;
    for (int sc : scenes) {
        for (int i : units_of[sc]) {
            while (fill[p] == size[p]) p = (p + 1) % parts;
            out[begin[p] + fill[p]++] = work[i];
        }
        p = (p + 1) % parts;
    }
}

int ensure_scratch(gudni_ctx *ctx)
{
    if (ctx->fast_blocks) return GUDNI_OK;
    ++ctx->scratch_gen;
    const int waves_per_block = 4;
    int per_cu = 0;
    CHECK_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dev::raster_pass, 256, 0));
    if (per_cu < 1) per_cu = 1;
    if (ctx->blocks_per_cu > 0) per_cu = std::min(per_cu, ctx->blocks_per_cu);
    ctx->fast_blocks = ctx->cus * per_cu;
    ctx->fast_blocks = std::min(ctx->fast_blocks, ctx->cus * 8);
    ctx->slow_blocks = std::max(1, ctx->cus / 8);
    const size_t fast_slots = (size_t)ctx->fast_blocks * waves_per_block;
    const size_t slow_slots = (size_t)ctx->slow_blocks * waves_per_block;
    const size_t fcap = (size_t)1 << dev::kFastCapLog2;
    int slog = 0;
    while ((1 << slog) < ctx->spec.max_thresholds) ++slog;
    const size_t scap = (size_t)1 << slog;
    CHECK_HIP(ctx, ctx->fast_thr.ensure(fast_slots * fcap * 64 * 16));
    CHECK_HIP(ctx, ctx->fast_hdr.ensure(fast_slots * fcap * 64 * 4));
    CHECK_HIP(ctx, ctx->fast_sidx.ensure(fast_slots * dev::kSidxWords * 64 * 4));
    CHECK_HIP(ctx, ctx->slow_thr.ensure(slow_slots * scap * 64 * 16));
    CHECK_HIP(ctx, ctx->slow_hdr.ensure(slow_slots * scap * 64 * 4));
    CHECK_HIP(ctx, ctx->slow_sidx.ensure(slow_slots * dev::kSidxWords * 64 * 4));
    return GUDNI_OK;
}

int ensure_control(gudni_ctx *ctx, int n_work)
{
    const int want = std::max(1024, n_work);
    if (ctx->retry_capacity < want) {
        ++ctx->scratch_gen;
        CHECK_HIP(ctx, ctx->control.ensure(sizeof(uint32_t) * (size_t)(dev::kCtlWords + want)));
        ctx->retry_capacity = want;
    }
    return GUDNI_OK;
}

// Fill a Frame with the context's buffers; caller sets the per-launch parts.
dev::Frame base_frame(gudni_ctx *ctx)
{
    dev::Frame f;
    std::memset(&f, 0, sizeof(f));
    f.random = (const float *)ctx->random.p;
    f.compute_depth = ceil_log2(ctx->spec.threads_per_tile);
    f.max_shapes = ctx->spec.max_shapes;
    f.max_thresholds = ctx->spec.max_thresholds;
    f.random_mask = ctx->spec.random_field_size - 1;
    f.stochastic = ctx->spec.stochastic_factor;
    f.coverage_mode = ctx->spec.coverage_mode;
    f.out_rgba = ctx->out_format == GUDNI_FORMAT_RGBA8 ? 1 : 0;
    f.threads_per_tile = ctx->spec.threads_per_tile;
    return f;
}

// The one-entry scene table of a single-frame launch; uploaded only when a
// field changed since the last launch (normally never between frames).
static int set_single_scene(gudni_ctx *ctx, const dev::SceneDesc &sd, dev::Frame &f)
{
    if (!ctx->scene_table.p || std::memcmp(&sd, &ctx->single_scene, sizeof(sd)) != 0) {
        CHECK_HIP(ctx, ctx->scene_table.ensure(sizeof(dev::SceneDesc)));
        ctx->single_scene = sd;
        CHECK_HIP(ctx, hipMemcpyAsync(ctx->scene_table.p, &ctx->single_scene, sizeof(sd),
                                      hipMemcpyHostToDevice, ctx->stream));
    }
    f.scenes = (const dev::SceneDesc *)ctx->scene_table.p;
    f.tile_scene = nullptr;
    f.scene_done = nullptr;
    return GUDNI_OK;
}

static dev::SceneDesc single_desc(gudni_ctx *ctx, const float bg[4], int32_t w, int32_t h,
                                  uint32_t *out, float *out_f32)
{
    dev::SceneDesc sd;
    std::memset(&sd, 0, sizeof(sd));
    sd.geo = (const uint8_t *)ctx->geo.p;
    sd.shapes = (const ShapeRec *)ctx->shapes.p;
    sd.bounds = (const float4 *)ctx->xrange.p;
    sd.substances = (const float4 *)ctx->subs.p;
    sd.pict = (const uint8_t *)ctx->pict.p;
    sd.uses = (const PictUse *)ctx->uses.p;
    sd.out = out;
    sd.out_f32 = (float4 *)out_f32;
    sd.grads = (const float4 *)ctx->grads.p;
    for (int i = 0; i < 4; ++i) sd.bg[i] = bg[i];
    sd.width = w;
    sd.height = h;
    sd.filter = ctx->spec.texture_filter;
    sd.sort_bits = ctx->spec.sort_radix_bits;
    return sd;
}

// Both passes, stream-ordered.  The retry pass runs unconditionally; with no
// abandoned units its waves exit on their first claim.
// Where the two passes keep their control words, rings and shape scratch,
// and their launch shapes.
struct PassBuffers {
    uint32_t *ctl;               // counters, status word, retry list
    int32_t retry_capacity;
    float4 *fast_thr;
    uint32_t *fast_hdr;
    int32_t *fast_sidx;
    int fast_blocks;             // blocks of 256 threads
    float4 *slow_thr;
    uint32_t *slow_hdr;
    int32_t *slow_sidx;
    int slow_blocks, slow_threads;
};

static PassBuffers ctx_buffers(gudni_ctx *ctx)
{
    return PassBuffers{(uint32_t *)ctx->control.p, ctx->retry_capacity,
                       (float4 *)ctx->fast_thr.p, (uint32_t *)ctx->fast_hdr.p, (int32_t *)ctx->fast_sidx.p,
                       ctx->fast_blocks,
                       (float4 *)ctx->slow_thr.p, (uint32_t *)ctx->slow_hdr.p, (int32_t *)ctx->slow_sidx.p,
                       ctx->slow_blocks, 256};
}

static int launch_passes(gudni_ctx *ctx, dev::Frame f, const PassBuffers &b)
{
    uint32_t *ctl = b.ctl;
    hipLaunchKernelGGL(dev::reset_counters, dim3(1), dim3(64), 0, ctx->stream, ctl, dev::kCtlWords);
    CHECK_HIP(ctx, hipGetLastError());
    f.status = ctl + dev::kCtlStatus;
    ctx->status_word = f.status;
    f.retry = ctl + dev::kCtlRetry;
    f.retry_capacity = b.retry_capacity;
    // fast pass: small rings, as many waves as fit
    f.counter = ctl + dev::kCtlFast;
    f.n_parts = dev::kParts;
    f.q_thr = b.fast_thr;
    f.q_hdr = b.fast_hdr;
    f.shape_idx = b.fast_sidx;
    f.cap_log2 = dev::kFastCapLog2;
    const int fast_grid = std::max(1, std::min(b.fast_blocks, (f.n_work + 3) / 4));
    hipLaunchKernelGGL(dev::raster_pass, dim3(fast_grid), dim3(256), 0, ctx->stream, f, 0);
    CHECK_HIP(ctx, hipGetLastError());
    if (ctx->timed_render) CHECK_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    // retry pass: rings of maxThresholds records
    int slog = 0;
    while ((1 << slog) < ctx->spec.max_thresholds) ++slog;
    f.counter = ctl + dev::kCtlSlow;
    f.n_parts = 1;
    f.q_thr = b.slow_thr;
    f.q_hdr = b.slow_hdr;
    f.shape_idx = b.slow_sidx;
    f.cap_log2 = slog;
    hipLaunchKernelGGL(dev::raster_pass, dim3(b.slow_blocks), dim3(b.slow_threads), 0, ctx->stream, f, 1);
    CHECK_HIP(ctx, hipGetLastError());
    return GUDNI_OK;
}

static int launch_passes(gudni_ctx *ctx, dev::Frame f) { return launch_passes(ctx, f, ctx_buffers(ctx)); }

int read_status(gudni_ctx *ctx)
{
    uint32_t st = 0;
    const uint32_t *word = ctx->status_word ? ctx->status_word : (uint32_t *)ctx->control.p + dev::kCtlStatus;
    CHECK_HIP(ctx, hipMemcpyAsync(&st, word, 4, hipMemcpyDeviceToHost, ctx->stream));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (st & dev::kStatusOverflow)
        return set_error(ctx, GUDNI_ERR_OVERFLOW, "a column's threshold queue exceeded maxThresholds");
    if (st & dev::kStatusGuard)
        return set_error(ctx, GUDNI_ERR_GUARD, "a column exceeded the section-count guard");
    return GUDNI_OK;
}

CommonView ctx_view(gudni_ctx *ctx)
{
    return CommonView{ctx->geo_host.data(), ctx->geo_host.size(), ctx->subs_host.data(),
                      ctx->n_substances, ctx->n_uses, ctx->n_grads};
}

int validate_gradients(gudni_ctx *ctx, const void *grads, int32_t n)
{
    if (n < 0 || (n && !grads)) return set_error(ctx, GUDNI_ERR_ARG, "bad gradient table");
    const gudni_gradient *g = (const gudni_gradient *)grads;
    for (int32_t i = 0; i < n; ++i)
        if ((g[i].kind != GUDNI_GRADIENT_LINEAR && g[i].kind != GUDNI_GRADIENT_RADIAL) ||
            g[i].n_stops < 1 || g[i].n_stops > 4)
            return set_error(ctx, GUDNI_ERR_ARG, "gradient " + std::to_string(i) +
                                                     ": kind must be 0 or 1 with 1..4 stops");
    return GUDNI_OK;
}

int validate_uses(gudni_ctx *ctx, const void *pict_uses, int32_t n_pict_uses,
                         size_t pict_bytes)
{
    for (int32_t i = 0; i < n_pict_uses; ++i) {
        PictUse u;
        std::memcpy(&u, (const uint8_t *)pict_uses + 24 * (size_t)i, 24);
        const int64_t need = (int64_t)u.mem_offset + 4 * (int64_t)u.w * (int64_t)u.h;
        if (u.w < 0 || u.h < 0 || u.mem_offset < 0 || (u.mem_offset & 3) || need > (int64_t)pict_bytes)
            return set_error(ctx, GUDNI_ERR_ARG, "picture usage " + std::to_string(i) +
                                                     " outside the picture data");
        // the kernels address texels with 32-bit byte offsets
        if (need > (int64_t)INT32_MAX)
            return set_error(ctx, GUDNI_ERR_ARG, "picture usage " + std::to_string(i) +
                                                     " reaches beyond 2 GiB of picture data");
    }
    return GUDNI_OK;
}

// Host-side validation of a job's records against the common buffers, so no
// kernel ever reads or writes outside them: tile slices inside the job's
// shapes, pixel origins non-negative, every referenced strand walk inside the
// geometry heap with even 8-byte sizes >= 4 (so float4 reads stay aligned),
// substance ids inside the substance table and picture ids inside the usage
// table.
// One shape reference of a job: its substance and strands must lie inside
// the uploaded tables and heap.  Writes the reference's extent (the
// generation prefilter's input) and returns null, or returns what is wrong.
static const char *ref_extent(const CommonView &cv, const ShapeRec &r, float *extent)
{
    const uint64_t sid = r.tag & dev::kTagIdMask;
    if (sid >= (uint64_t)cv.n_subs) return "shape substance id outside the substance table";
    const uint64_t type = r.tag & dev::kTagTypeMask;
    if (type != dev::kTagSolid) {
        const uint32_t id = __builtin_bit_cast(uint32_t, cv.subs[sid].x);
        if (type == dev::kTagGradient) {
            if (id >= (uint32_t)cv.n_grads) return "gradient substance refers to a missing gradient";
        } else if (id >= (uint32_t)cv.n_uses) {
            return "picture substance refers to a missing usage";
        }
    }
    const size_t geo_bytes = cv.geo_bytes;
    size_t pos = 16u * (size_t)r.geo_start;
    // Extent of the shape's strands: x from the end points, y over every
    // point (curves stay inside their control hull).  NaN coordinates can
    // never pass the kernel's range tests, so they are left out; a NaN y
    // makes the y extent unknown, which disables the y test.
    float xlo = INFINITY, xhi = -INFINITY, ylo = INFINITY, yhi = -INFINITY;
    bool y_nan = false;
    for (uint32_t k = 0; k < r.num_strands; ++k) {
        if (pos + 8 > geo_bytes) return "strand header outside the geometry heap";
        uint16_t size;
        std::memcpy(&size, cv.geo + pos, 2);
        if (size < 4 || (size & 1) || pos + 8u * size > geo_bytes) return "malformed strand in the geometry heap";
        const float *pts = reinterpret_cast<const float *>(cv.geo + pos + 8);
        const int npts = (int)size - 1;         // right, left, control, tree pairs
        if (pts[2] < xlo) xlo = pts[2];
        if (pts[0] > xhi) xhi = pts[0];
        for (int p = 0; p < npts; ++p) {
            const float y = pts[2 * p + 1];
            if (y != y) y_nan = true;
            if (y < ylo) ylo = y;
            if (y > yhi) yhi = y;
        }
        pos += 8u * size;
    }
    if (y_nan) ylo = -INFINITY;
    extent[0] = xlo;
    extent[1] = ylo;
    extent[2] = xhi;
    extent[3] = yhi;
    return nullptr;
}

int validate_job(gudni_ctx *ctx, const CommonView &cv, const ShapeRec *shapes,
                        int32_t n_shapes, const TileRec *tiles, int32_t n_tiles, int32_t w,
                        int32_t h, std::vector<float> &xrange)
{
    if (w <= 0 || h <= 0 || n_tiles < 0 || n_shapes < 0)
        return set_error(ctx, GUDNI_ERR_ARG, "bad job size");
    for (int32_t t = 0; t < n_tiles; ++t) {
        const TileRec &r = tiles[t];
        if (r.shape_start < 0 || r.shape_count < 0 || (int64_t)r.shape_start + r.shape_count > n_shapes)
            return set_error(ctx, GUDNI_ERR_ARG, "tile shape slice outside the job's shapes");
        if (r.hdepth < 0 || r.hdepth > 8 || r.vdepth < 0 || r.vdepth > 12)
            return set_error(ctx, GUDNI_ERR_ARG, "tile depth out of range");
        if (r.box[0] < 0 || r.box[1] < 0)
            return set_error(ctx, GUDNI_ERR_ARG, "tile box has a negative origin");
    }
    // the shape references in parallel chunks; the first malformed reference
    // (lowest index) is the one reported
    const int64_t n = n_shapes;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = n >= 65536 ? std::min(16u, hw) : 1u;
    std::vector<int64_t> bad(nt, -1);
    std::vector<const char *> why(nt, nullptr);
    auto chunk = [&](unsigned t) {
        const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        for (int64_t i = lo; i < hi; ++i) {
            const char *m = ref_extent(cv, shapes[i], &xrange[4 * (size_t)i]);
            if (m) {
                bad[t] = i;
                why[t] = m;
                return;
            }
        }
    };
    std::vector<std::thread> pool;
    unsigned started = 1;   // chunks 1 .. started-1 have a thread of their own
    try {
        for (unsigned t = 1; t < nt; ++t) {
            pool.emplace_back(chunk, t);
            ++started;
        }
    } catch (const std::system_error &) {
        // no more threads: the calling thread takes the remaining chunks
    }
    chunk(0);
    for (unsigned t = started; t < nt; ++t) chunk(t);
    for (std::thread &th : pool) th.join();
    for (unsigned t = 0; t < nt; ++t)
        if (bad[t] >= 0) return set_error(ctx, GUDNI_ERR_ARG, why[t]);
    return GUDNI_OK;
}

int bin_scene(gudni_ctx *ctx, const BinIn &in, const BinOut &out, BinCounts &cnt)
{
    binning::Params bp;
    bp.n_shapes = in.n_shapes;
    bp.canvas_depth = ceil_log2(std::max(in.width, in.height));
    bp.tile_depth = ceil_log2(in.max_tile_size > 0 ? in.max_tile_size : ctx->spec.max_tile_size);
    bp.cap_shapes = ctx->spec.max_shapes - 1;
    bp.cap_strands = (uint32_t)ctx->spec.max_strands_per_tile;
    bp.per_job = ctx->spec.threads_per_tile;
    bp.col_step = ctx->spec.max_tiles_per_call;
    bp.threads_per_tile = ctx->spec.threads_per_tile;
    bp.width = in.width;
    bp.height = in.height;
    // The root is a 2^cd square and the levels halve height and width in
    // turn; a node whose side in its split direction is 8 or less is a leaf,
    // so the deepest level is 2 (cd - 3).  One level more for margin (a
    // deeper tree stops with an error, never silently).
    bp.max_levels = std::max(2, 2 * (bp.canvas_depth - 3) + 2);
    if (!ctx->bin_cap_pairs) {
        ctx->bin_cap_pairs = (int32_t)std::min<int64_t>(std::max<int64_t>(1 << 16, 8 * (int64_t)bp.n_shapes), 1 << 28);
        ctx->bin_cap_nodes = std::max(1 << 15, bp.n_shapes / 4);
    }
    if (ctx->timing)
        for (hipEvent_t &e : ctx->bin_ev)
            if (!e) CHECK_HIP(ctx, hipEventCreate(&e));
    uint32_t status[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int attempt = 0;; ++attempt) {
        const int32_t cp = std::max(ctx->bin_cap_pairs, bp.n_shapes), cn = ctx->bin_cap_nodes;
        CHECK_HIP(ctx, ctx->bin_ws.ensure(binning::workspace_bytes(cp, cn, bp.max_levels)));
        CHECK_HIP(ctx, out.tiles->ensure(32u * (size_t)cn));
        CHECK_HIP(ctx, out.tile_job->ensure(4u * (size_t)cn));
        CHECK_HIP(ctx, out.refs->ensure(16u * (size_t)cp));
        CHECK_HIP(ctx, out.ref_bounds->ensure(16u * (size_t)cp));
        CHECK_HIP(ctx, out.work->ensure(16u * (size_t)cn));
        binning::Workspace ws;
        binning::carve(ws, ctx->bin_ws.p, cp, cn, bp.max_levels);
        binning::Output o{(TileRec *)out.tiles->p, (int32_t *)out.tile_job->p, (ShapeRec *)out.refs->p,
                          (float4 *)out.ref_bounds->p, (uint32_t *)out.work->p};
        if (ctx->timing) CHECK_HIP(ctx, hipEventRecord(ctx->bin_ev[0], ctx->stream));
        CHECK_HIP(ctx, binning::enqueue(bp, ws, in.shapes, in.bbox, in.bounds, o, ctx->stream));
        if (ctx->timing) CHECK_HIP(ctx, hipEventRecord(ctx->bin_ev[1], ctx->stream));
        CHECK_HIP(ctx, hipMemcpyAsync(status, ws.status, sizeof(status), hipMemcpyDeviceToHost,
                                      ctx->stream));
        CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (status[0] == 0) break;
        if (status[0] != 1)
            return set_error(ctx, GUDNI_ERR_ARG, "scene binning: tile tree deeper than the canvas allows");
        // the capacity that ran out grows 4x, up to 2^28 entries (32-bit
        // indices throughout); one already at the clamp fails at once
        constexpr int32_t kCapMax = 1 << 28;
        const bool over_nodes = (status[6] & 1u) != 0, over_pairs = (status[6] & 2u) != 0;
        if (attempt >= 8 || (over_nodes && cn >= kCapMax) || (over_pairs && cp >= kCapMax))
            return set_error(ctx, GUDNI_ERR_ARG, over_nodes && cn >= kCapMax
                                 ? "scene binning: tile tree needs more than 2^28 nodes"
                                 : "scene binning: tile tree needs more than 2^28 (node, shape) references");
        if (over_pairs) ctx->bin_cap_pairs = (int32_t)std::min<int64_t>(4 * (int64_t)cp, kCapMax);
        if (over_nodes) ctx->bin_cap_nodes = (int32_t)std::min<int64_t>(4 * (int64_t)cn, kCapMax);
    }
    if (ctx->timing) CHECK_HIP(ctx, hipEventElapsedTime(&ctx->last_bin_ms, ctx->bin_ev[0], ctx->bin_ev[1]));
    cnt.n_tiles = (int32_t)status[1];
    cnt.n_refs = (int32_t)status[2];
    cnt.n_work = (int32_t)status[3];
    return GUDNI_OK;
}

int bin_scenes(gudni_ctx *ctx, const BinIn *in, const BinOut *out, BinCounts *cnt, int n)
{
    if (n <= 0) return GUDNI_OK;
    std::vector<binning::Params> params((size_t)n);
    std::vector<int32_t> cap_pairs((size_t)n), cap_nodes((size_t)n), max_nodes((size_t)n);
    for (int s = 0; s < n; ++s) {
        binning::Params &bp = params[s];
        bp.n_shapes = in[s].n_shapes;
        bp.canvas_depth = ceil_log2(std::max(in[s].width, in[s].height));
        bp.tile_depth = ceil_log2(in[s].max_tile_size > 0 ? in[s].max_tile_size : ctx->spec.max_tile_size);
        bp.cap_shapes = ctx->spec.max_shapes - 1;
        bp.cap_strands = (uint32_t)ctx->spec.max_strands_per_tile;
        bp.per_job = ctx->spec.threads_per_tile;
        bp.col_step = ctx->spec.max_tiles_per_call;
        bp.threads_per_tile = ctx->spec.threads_per_tile;
        bp.width = in[s].width;
        bp.height = in[s].height;
        bp.max_levels = std::max(2, 2 * (bp.canvas_depth - 3) + 2);
        // a tree down to the 8-pixel floor has at most 2 (side / 8)^2 nodes;
        // start below that for large canvases and grow on overflow
        const int64_t cells = (int64_t)std::max(1, (1 << bp.canvas_depth) / 8);
        max_nodes[s] = (int32_t)std::min<int64_t>(2 * cells * cells + 16, 1 << 28);
        cap_nodes[s] = (int32_t)std::min<int64_t>(max_nodes[s], std::max<int64_t>(1 << 15, 4 * (int64_t)bp.n_shapes));
        // the work list's scan reuses the pair scan's words: room for a
        // count per tile, and tiles never outnumber nodes
        cap_pairs[s] = (int32_t)std::max<int64_t>(
            std::min<int64_t>(std::max<int64_t>(1 << 14, 8 * (int64_t)bp.n_shapes), 1 << 28), cap_nodes[s]);
    }
    std::vector<int> todo((size_t)n);
    for (int s = 0; s < n; ++s) todo[s] = s;
    for (int attempt = 0; !todo.empty(); ++attempt) {
        if (attempt > 8) return set_error(ctx, GUDNI_ERR_ARG, "scene binning: capacities keep overflowing");
        const size_t m = todo.size();
        // workspaces after one status array of 8 words per scene
        std::vector<size_t> off(m);
        size_t total = (32 * m + 255) & ~(size_t)255;
        for (size_t i = 0; i < m; ++i) {
            const int s = todo[i];
            off[i] = total;
            total += binning::workspace_bytes(cap_pairs[s], cap_nodes[s], params[s].max_levels);
        }
        CHECK_HIP(ctx, ctx->bins_ws.ensure(total));
        std::vector<binning::SceneJob> jobs(m);
        for (size_t i = 0; i < m; ++i) {
            const int s = todo[i];
            const BinOut &o = out[s];
            CHECK_HIP(ctx, o.tiles->ensure(32u * (size_t)cap_nodes[s]));
            CHECK_HIP(ctx, o.tile_job->ensure(4u * (size_t)cap_nodes[s]));
            CHECK_HIP(ctx, o.refs->ensure(16u * (size_t)cap_pairs[s]));
            CHECK_HIP(ctx, o.ref_bounds->ensure(16u * (size_t)cap_pairs[s]));
            CHECK_HIP(ctx, o.work->ensure(16u * (size_t)cap_nodes[s]));
            binning::SceneJob &j = jobs[i];
            j.p = params[s];
            binning::carve(j.w, (uint8_t *)ctx->bins_ws.p + off[i], cap_pairs[s], cap_nodes[s], params[s].max_levels);
            j.w.status = (uint32_t *)ctx->bins_ws.p + 8 * i;
            j.shapes = in[s].shapes;
            j.bbox = in[s].bbox;
            j.shape_bounds = in[s].bounds;
            j.out = binning::Output{(TileRec *)o.tiles->p, (int32_t *)o.tile_job->p, (ShapeRec *)o.refs->p,
                                    (float4 *)o.ref_bounds->p, (uint32_t *)o.work->p};
        }
        CHECK_HIP(ctx, upload(ctx, ctx->bins_jobs, jobs.data(), sizeof(binning::SceneJob) * m));
        CHECK_HIP(ctx, binning::enqueue_scenes((const binning::SceneJob *)ctx->bins_jobs.p, (int)m, ctx->stream));
        std::vector<uint32_t> status(8 * m);
        CHECK_HIP(ctx, hipMemcpyAsync(status.data(), ctx->bins_ws.p, 4 * status.size(), hipMemcpyDeviceToHost,
                                      ctx->stream));
        CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        std::vector<int> again;
        for (size_t i = 0; i < m; ++i) {
            const int s = todo[i];
            const uint32_t *st = status.data() + 8 * i;
            if (st[0] == 0) {
                cnt[s].n_tiles = (int32_t)st[1];
                cnt[s].n_refs = (int32_t)st[2];
                cnt[s].n_work = (int32_t)st[3];
            } else if (st[0] == 1) {
                // grow only what ran out; fail once it is at its clamp
                const bool over_nodes = (st[6] & 1u) != 0, over_pairs = (st[6] & 2u) != 0;
                if ((over_nodes && cap_nodes[s] >= max_nodes[s]) || (over_pairs && cap_pairs[s] >= (1 << 28)))
                    return set_error(ctx, GUDNI_ERR_ARG, "scene binning: tile tree exceeds the node or reference limit");
                if (over_nodes) cap_nodes[s] = (int32_t)std::min<int64_t>(4 * (int64_t)cap_nodes[s], max_nodes[s]);
                if (over_pairs) cap_pairs[s] = (int32_t)std::min<int64_t>(4 * (int64_t)cap_pairs[s], 1 << 28);
                cap_pairs[s] = std::max(cap_nodes[s], cap_pairs[s]);
                again.push_back(s);
            } else {
                return set_error(ctx, GUDNI_ERR_ARG, "scene binning: tile tree deeper than the canvas allows");
            }
        }
        todo.swap(again);
    }
    return GUDNI_OK;
}

} // namespace gudni

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Appends `bytes` of device memory at an `align`-ed offset of a growing
// device array (capacity doubles; the used prefix moves along).
hipError_t append_dev(DevBuf &buf, size_t &used, const void *src, size_t bytes, size_t align, hipStream_t st,
                      size_t &offset)
{
    const size_t at = align_up(used, align), need = at + bytes;
    if (need > buf.bytes || !buf.p) {
        const size_t cap = std::max(std::max(need, 2 * buf.bytes), (size_t)256);
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, cap);
        if (e != hipSuccess) return e;
        if (used && buf.p) {
            e = hipMemcpyAsync(p, buf.p, used, hipMemcpyDeviceToDevice, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) { (void)hipFree(p); return e; }
        }
        if (buf.p) (void)hipFree(buf.p);
        buf.p = p;
        buf.bytes = cap;
    }
    offset = at;
    used = need;
    return bytes ? hipMemcpyAsync((uint8_t *)buf.p + at, src, bytes, hipMemcpyDeviceToDevice, st) : hipSuccess;
}

__global__ void rebase_work(uint32_t *work, int n, uint32_t tile_base)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        work[i] += tile_base << 2;
}
__global__ void fill_i32(int32_t *dst, int n, int32_t v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = v;
}

}  // namespace

extern "C" {

int gudni_create(const gudni_spec *spec, int device, gudni_ctx **out)
{
    if (!out) return GUDNI_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) {
        g_host_error = "gudni_create: no HIP device";
        return GUDNI_ERR_NO_DEVICE;
    }
    gudni_ctx *ctx = new gudni_ctx();
    if (spec) ctx->spec = *spec;
    else gudni_default_spec(&ctx->spec);
    const gudni_spec &s = ctx->spec;
    if (s.threads_per_tile != 256 || s.max_shapes < 1 || s.max_shapes > 127 || s.max_tile_size < 8 ||
        s.max_tile_size > s.threads_per_tile || (s.max_tile_size & (s.max_tile_size - 1)) ||
        s.max_thresholds < 1 || s.random_field_size < 1 ||
        (s.coverage_mode != GUDNI_COVERAGE_ACCURATE && s.coverage_mode != GUDNI_COVERAGE_REFERENCE) ||
        (s.texture_filter != GUDNI_FILTER_NEAREST && s.texture_filter != GUDNI_FILTER_BILINEAR) ||
        s.sort_radix_bits < 0 || s.sort_radix_bits > dev::kRadixMaxBits ||
        (s.random_field_size & (s.random_field_size - 1))) {
        g_host_error = "gudni_create: unsupported spec (threads_per_tile must be 256, "
                       "max_tile_size a power of two in 8..256, "
                       "max_shapes 1..127, random_field_size a power of two, "
                       "sort_radix_bits 0..5)";
        delete ctx;
        return GUDNI_ERR_ARG;
    }
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete ctx;
        return GUDNI_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete ctx;
        return GUDNI_ERR_HIP;
    }
    ctx->cus = prop.multiProcessorCount;
    // the static scratch's address on this device (for
    // gudni_frame_render_static; a failure surfaces only there)
    lookup_static_scratch(device);
    *out = ctx;
    return GUDNI_OK;
}

void gudni_destroy(gudni_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    DevBuf *bufs[] = {&ctx->geo, &ctx->subs, &ctx->pict, &ctx->uses, &ctx->random, &ctx->grads,
                      &ctx->tiles, &ctx->tile_job, &ctx->shapes, &ctx->xrange, &ctx->work, &ctx->out,
                      &ctx->out_f32, &ctx->fast_thr, &ctx->fast_hdr, &ctx->fast_sidx,
                      &ctx->slow_thr, &ctx->slow_hdr, &ctx->slow_sidx, &ctx->control, &ctx->dump,
                      &ctx->cycles, &ctx->scene_table, &ctx->scene_shapes, &ctx->scene_bbox,
                      &ctx->scene_bounds, &ctx->bin_ws, &ctx->ser_tags, &ctx->ser_soff,
                      &ctx->ser_ooff, &ctx->ser_pairs, &ctx->ser_work, &ctx->bins_ws, &ctx->bins_jobs,
                      &ctx->ser_many_ws, &ctx->ser_many_args};
    for (DevBuf *b : bufs) b->release();
    for (hipEvent_t &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t &e : ctx->bin_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t &e : ctx->ser_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->l1_stage) (void)hipHostFree(ctx->l1_stage);
    delete ctx;
}

const char *gudni_last_error(const gudni_ctx *ctx)
{
    return ctx ? ctx->error.c_str() : g_host_error.c_str();
}

int gudni_set_stream(gudni_ctx *ctx, void *stream)
{
    if (!ctx) return GUDNI_ERR_ARG;
    ctx->stream = (hipStream_t)stream;
    return GUDNI_OK;
}

int gudni_upload_common(gudni_ctx *ctx, const void *geo, size_t geo_bytes, const void *substances,
                        int32_t n_substances, const void *pict_data, size_t pict_bytes,
                        const void *pict_uses, int32_t n_pict_uses, const float *random_field,
                        int32_t n_random)
{
    if (!ctx) return GUDNI_ERR_ARG;
    if (n_substances < 0 || n_pict_uses < 0 || !random_field ||
        n_random != ctx->spec.random_field_size)
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_upload_common: bad sizes (random field "
                                             "must hold random_field_size floats)");
    if ((geo_bytes && !geo) || (n_substances && !substances) || (pict_bytes && !pict_data) ||
        (n_pict_uses && !pict_uses))
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_upload_common: null buffer");
    {
        const int rc = validate_uses(ctx, pict_uses, n_pict_uses, pict_bytes);
        if (rc) return rc;
    }
    ctx->geo_host.assign((const uint8_t *)geo, (const uint8_t *)geo + geo_bytes);
    ctx->subs_host.resize((size_t)n_substances);
    if (n_substances) std::memcpy(ctx->subs_host.data(), substances, 16u * (size_t)n_substances);
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, upload(ctx, ctx->geo, (const uint8_t *)geo, geo_bytes));
    CHECK_HIP(ctx, upload(ctx, ctx->subs, (const uint8_t *)substances, 16u * (size_t)n_substances));
    CHECK_HIP(ctx, upload(ctx, ctx->pict, (const uint8_t *)pict_data, pict_bytes));
    CHECK_HIP(ctx, upload(ctx, ctx->uses, (const uint8_t *)pict_uses, 24u * (size_t)n_pict_uses));
    CHECK_HIP(ctx, upload(ctx, ctx->random, random_field, 4u * (size_t)n_random));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->n_substances = n_substances;
    ctx->n_uses = n_pict_uses;
    ctx->frame_ready = false;
    return GUDNI_OK;
}

int gudni_upload_gradients(gudni_ctx *ctx, const void *gradients, int32_t n_gradients)
{
    if (!ctx) return GUDNI_ERR_ARG;
    int rc = validate_gradients(ctx, gradients, n_gradients);
    if (rc) return rc;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, upload(ctx, ctx->grads, (const uint8_t *)gradients, sizeof(gudni_gradient) * (size_t)n_gradients));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->n_grads = n_gradients;
    ctx->frame_ready = false;
    return GUDNI_OK;
}

static int run_job(gudni_ctx *ctx, const void *shapes, int32_t n_shapes, const void *tiles,
                   int32_t n_tiles, const float bg[4], int32_t w, int32_t h, int32_t frame,
                   int32_t job_index, uint32_t *host_out, gudni_dump *dump)
{
    (void)frame;
    if (!ctx) return GUDNI_ERR_ARG;
    if (!ctx->random.p) return set_error(ctx, GUDNI_ERR_ARG, "gudni_upload_common was not called");
    std::vector<float> xr(4 * (size_t)n_shapes + 4);
    int rc = validate_job(ctx, ctx_view(ctx), (const ShapeRec *)shapes, n_shapes,
                          (const TileRec *)tiles, n_tiles, w, h, xr);
    if (rc) return rc;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    // a job's records go into the buffers an uploaded frame (or binned
    // scene) renders from: that frame must be uploaded or binned again
    ctx->frame_ready = false;
    rc = ensure_scratch(ctx);
    if (rc) return rc;
    std::vector<uint32_t> work;
    build_work(ctx->spec, (const TileRec *)tiles, n_tiles, w, h, work);
    rc = ensure_control(ctx, (int)work.size());
    if (rc) return rc;
    std::vector<int32_t> tile_job((size_t)n_tiles, job_index);
    CHECK_HIP(ctx, upload(ctx, ctx->tiles, (const uint8_t *)tiles, 32u * (size_t)n_tiles));
    CHECK_HIP(ctx, upload(ctx, ctx->tile_job, tile_job.data(), 4u * tile_job.size()));
    CHECK_HIP(ctx, upload(ctx, ctx->shapes, (const uint8_t *)shapes, 16u * (size_t)n_shapes));
    CHECK_HIP(ctx, upload(ctx, ctx->xrange, xr.data(), 16u * (size_t)n_shapes));
    CHECK_HIP(ctx, upload(ctx, ctx->work, work.data(), 4u * work.size()));
    const size_t npix = (size_t)w * (size_t)h;
    CHECK_HIP(ctx, ctx->out.ensure(4 * npix));
    // the pixels a job can write lie in the box of its tiles' columns: only
    // that box of the host frame travels, both ways, so every other pixel
    // stays as it was.  A column is active while its row origin lies inside
    // the box and the canvas, and it writes 2^dd rows from there, clipped to
    // the canvas but not to the box (initTileState / isActiveThread,
    // Kernels.cl:1692-1722): a box shorter than its depths say is written
    // below its bottom.  Origins are >= 0.  The box goes through a pinned
    // staging buffer, row by row on the host, in one transfer each way.
    const int cd = ceil_log2(ctx->spec.threads_per_tile);
    int32_t by0 = h, by1 = 0, bx0 = w, bx1 = 0;
    for (int32_t t = 0; t < n_tiles; ++t) {
        const TileRec &r = ((const TileRec *)tiles)[t];
        const int64_t bh = (int64_t)r.box[3] - r.box[1];
        if (r.box[0] >= w || r.box[1] >= h || bh <= 0) continue;
        const int dd = std::max(0, (int)r.vdepth - (cd - (int)r.hdepth));
        const int64_t seg = (int64_t)1 << dd;
        const int64_t rows = std::min<int64_t>((bh + seg - 1) / seg, ctx->spec.threads_per_tile >> r.hdepth);
        by0 = std::min(by0, r.box[1]);
        by1 = (int32_t)std::max<int64_t>(by1, std::min<int64_t>(h, r.box[1] + rows * seg));
        bx0 = std::min(bx0, r.box[0]);
        bx1 = (int32_t)std::max<int64_t>(bx1, std::min<int64_t>(w, (int64_t)r.box[0] + ((int64_t)1 << r.hdepth)));
    }
    const bool window = host_out && by0 < by1 && bx0 < bx1;
    const size_t bw = window ? (size_t)(bx1 - bx0) : 0, bh = window ? (size_t)(by1 - by0) : 0;
    uint32_t *box_dev = (uint32_t *)ctx->out.p + (size_t)by0 * (size_t)w + (size_t)bx0;
    if (window) {
        // a call that failed after queueing its copy from the staging buffer
        // returned without waiting for it: no copy may still read the buffer
        // that is refilled (or freed) here
        CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->l1_stage_bytes < 4 * bw * bh) {
            if (ctx->l1_stage) CHECK_HIP(ctx, hipHostFree(ctx->l1_stage));
            ctx->l1_stage = nullptr;
            ctx->l1_stage_bytes = 0;
            CHECK_HIP(ctx, hipHostMalloc(&ctx->l1_stage, 4 * bw * bh, hipHostMallocDefault));
            ctx->l1_stage_bytes = 4 * bw * bh;
        }
        uint32_t *stage = (uint32_t *)ctx->l1_stage;
        for (size_t y = 0; y < bh; ++y)
            std::memcpy(stage + y * bw, host_out + (size_t)(by0 + y) * w + bx0, 4 * bw);
        CHECK_HIP(ctx, hipMemcpy2DAsync(box_dev, 4 * (size_t)w, stage, 4 * bw, 4 * bw, bh, hipMemcpyHostToDevice,
                                        ctx->stream));
    }
    dev::Frame f = base_frame(ctx);
    f.tiles = (const TileRec *)ctx->tiles.p;
    f.tile_job = (const int32_t *)ctx->tile_job.p;
    f.work = (const uint32_t *)ctx->work.p;
    f.n_work = (int32_t)work.size();
    rc = set_single_scene(ctx, single_desc(ctx, bg, w, h, (uint32_t *)ctx->out.p, nullptr), f);
    if (rc) return rc;
    size_t slots = 0, o_pre = 0, o_preh = 0, o_thr = 0, o_hdr = 0, o_q = 0, o_bits = 0, o_stack = 0,
           o_sidx = 0, total = 0;
    if (dump) {
        slots = (size_t)n_tiles * ctx->spec.threads_per_tile;
        const size_t cap = (size_t)std::max(dump->cap, 1);
        const size_t ms = (size_t)ctx->spec.max_shapes;
        o_pre = 0;
        o_preh = o_pre + slots * cap * 16;
        o_thr = o_preh + slots * cap * 4;
        o_hdr = o_thr + slots * cap * 16;
        o_q = o_hdr + slots * cap * 4;
        o_bits = o_q + slots * 4;
        o_stack = (o_bits + slots * 4 + 7) & ~(size_t)7;
        o_sidx = o_stack + slots * 16;
        total = o_sidx + slots * ms * 4;
        CHECK_HIP(ctx, ctx->dump.ensure(total));
        CHECK_HIP(ctx, hipMemsetAsync(ctx->dump.p, 0, total, ctx->stream));
        // columns of quarters that hold no active column are never visited:
        // they read as inactive (-1), like every inactive column
        CHECK_HIP(ctx, hipMemsetAsync((uint8_t *)ctx->dump.p + o_q, 0xFF, slots * 4, ctx->stream));
        uint8_t *b = (uint8_t *)ctx->dump.p;
        f.dump_pre_thr = (float4 *)(b + o_pre);
        f.dump_pre_hdr = (uint32_t *)(b + o_preh);
        f.dump_thr = (float4 *)(b + o_thr);
        f.dump_hdr = (uint32_t *)(b + o_hdr);
        f.dump_qlen = (int32_t *)(b + o_q);
        f.dump_bits = (uint32_t *)(b + o_bits);
        f.dump_stack = (uint64_t *)(b + o_stack);
        f.dump_sidx = (int32_t *)(b + o_sidx);
        f.dump_cap = (int32_t)cap;
    }
    ctx->timed_render = false;
    rc = launch_passes(ctx, f);
    if (rc) return rc;
    rc = read_status(ctx);
    if (rc) return rc;
    if (window) {
        uint32_t *stage = (uint32_t *)ctx->l1_stage;
        CHECK_HIP(ctx, hipMemcpy2DAsync(stage, 4 * bw, box_dev, 4 * (size_t)w, 4 * bw, bh, hipMemcpyDeviceToHost,
                                        ctx->stream));
        CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (size_t y = 0; y < bh; ++y)
            std::memcpy(host_out + (size_t)(by0 + y) * w + bx0, stage + y * bw, 4 * bw);
    }
    if (dump) {
        std::vector<uint8_t> hb(total);
        CHECK_HIP(ctx, hipMemcpyAsync(hb.data(), ctx->dump.p, total, hipMemcpyDeviceToHost, ctx->stream));
        CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const size_t cap = (size_t)dump->cap;
        std::memcpy(dump->pre_thresholds, hb.data() + o_pre, slots * cap * 16);
        std::memcpy(dump->pre_headers, hb.data() + o_preh, slots * cap * 4);
        std::memcpy(dump->thresholds, hb.data() + o_thr, slots * cap * 16);
        std::memcpy(dump->headers, hb.data() + o_hdr, slots * cap * 4);
        std::memcpy(dump->queue_len, hb.data() + o_q, slots * 4);
        std::memcpy(dump->shape_bits, hb.data() + o_bits, slots * 4);
        std::memcpy(dump->shape_stack, hb.data() + o_stack, slots * 16);
        std::memcpy(dump->shape_indices, hb.data() + o_sidx, slots * (size_t)ctx->spec.max_shapes * 4);
    }
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GUDNI_OK;
}

int gudni_raster_job(gudni_ctx *ctx, const void *shapes, int32_t n_shapes, const void *tiles,
                     int32_t n_tiles, const float background[4], int32_t width, int32_t height,
                     int32_t frame, int32_t job_index, uint32_t *out)
{
    if (!ctx || !out || !background || (n_tiles && !tiles) || (n_shapes && !shapes))
        return ctx ? set_error(ctx, GUDNI_ERR_ARG, "gudni_raster_job: null argument") : GUDNI_ERR_ARG;
    return run_job(ctx, shapes, n_shapes, tiles, n_tiles, background, width, height, frame,
                   job_index, out, nullptr);
}

int gudni_dump_job(gudni_ctx *ctx, const void *shapes, int32_t n_shapes, const void *tiles,
                   int32_t n_tiles, int32_t width, int32_t height, int32_t job_index,
                   gudni_dump *dump)
{
    if (!ctx || !dump || dump->cap < 1) return GUDNI_ERR_ARG;
    const float bg[4] = {1, 1, 1, 1};
    return run_job(ctx, shapes, n_shapes, tiles, n_tiles, bg, width, height, 0, job_index,
                   nullptr, dump);
}

int gudni_frame_upload(gudni_ctx *ctx, const gudni_jobs *jobs, int32_t width, int32_t height,
                       const float background[4], int32_t frame)
{
    if (!ctx || !jobs || !background) return GUDNI_ERR_ARG;
    if (!ctx->random.p) return set_error(ctx, GUDNI_ERR_ARG, "gudni_upload_common was not called");
    gudni_jobs_view v;
    gudni_jobs_view_get(jobs, &v);
    const TileRec *src_tiles = (const TileRec *)v.tiles;
    std::vector<TileRec> tiles(src_tiles, src_tiles + v.n_tiles);
    std::vector<int32_t> tile_job((size_t)v.n_tiles);
    for (int j = 0; j < v.n_jobs; ++j) {
        for (int t = v.job_tile_offset[j]; t < v.job_tile_offset[j + 1]; ++t) {
            tiles[t].shape_start += v.job_shape_offset[j];   // absolute into all refs
            // queueRasterJobs numbers the reversed job list (CallKernels.hs:241, 255)
            tile_job[t] = v.n_jobs - 1 - j;
        }
    }
    std::vector<float> xr(4 * (size_t)v.n_refs + 4);
    int rc = validate_job(ctx, ctx_view(ctx), (const ShapeRec *)v.shape_refs, v.n_refs,
                          tiles.data(), v.n_tiles, width, height, xr);
    if (rc) return rc;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    rc = ensure_scratch(ctx);
    if (rc) return rc;
    std::vector<uint32_t> work;
    build_work(ctx->spec, tiles.data(), v.n_tiles, width, height, work);
    rc = ensure_control(ctx, (int)work.size());
    if (rc) return rc;
    CHECK_HIP(ctx, upload(ctx, ctx->tiles, tiles.data(), 32u * tiles.size()));
    CHECK_HIP(ctx, upload(ctx, ctx->tile_job, tile_job.data(), 4u * tile_job.size()));
    CHECK_HIP(ctx, upload(ctx, ctx->shapes, v.shape_refs, 16u * (size_t)v.n_refs));
    CHECK_HIP(ctx, upload(ctx, ctx->xrange, xr.data(), 16u * (size_t)v.n_refs));
    CHECK_HIP(ctx, upload(ctx, ctx->work, work.data(), 4u * work.size()));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->frame_w = width;
    ctx->frame_h = height;
    ctx->frame_n_work = (int32_t)work.size();
    ctx->frame_n_tiles = v.n_tiles;
    ctx->frame_index = frame;
    ctx->frame_sort_bits = jobs->j.sort_radix_bits;
    std::memcpy(ctx->background, background, sizeof(ctx->background));
    ctx->frame_ready = true;
    return GUDNI_OK;
}

int gudni_frame_render(gudni_ctx *ctx, uint32_t *out_dev, float *out_f32_dev)
{
    if (!ctx || !out_dev) return GUDNI_ERR_ARG;
    if (!ctx->frame_ready) return set_error(ctx, GUDNI_ERR_ARG, "gudni_frame_upload was not called");
    dev::Frame f = base_frame(ctx);
    f.tiles = (const TileRec *)ctx->tiles.p;
    f.tile_job = (const int32_t *)ctx->tile_job.p;
    f.work = (const uint32_t *)ctx->work.p;
    f.n_work = ctx->frame_n_work;
    dev::SceneDesc sd = single_desc(ctx, ctx->background, ctx->frame_w, ctx->frame_h, out_dev, out_f32_dev);
    sd.sort_bits = ctx->frame_sort_bits;
    int rc0 = set_single_scene(ctx, sd, f);
    if (rc0) return rc0;
    if (ctx->timing) {
        for (hipEvent_t &e : ctx->ev)
            if (!e) CHECK_HIP(ctx, hipEventCreate(&e));
        CHECK_HIP(ctx, ctx->cycles.ensure(8 * sizeof(unsigned long long)));
        hipLaunchKernelGGL(dev::reset_counters, dim3(1), dim3(64), 0, ctx->stream,
                           (uint32_t *)ctx->cycles.p, 16);
        CHECK_HIP(ctx, hipGetLastError());
        f.stage_cycles = (unsigned long long *)ctx->cycles.p;
        CHECK_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    }
    ctx->timed_render = ctx->timing;
    int rc = launch_passes(ctx, f);
    if (rc) return rc;
    if (ctx->timing) CHECK_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    return GUDNI_OK;
}

int gudni_frame_render_static(gudni_ctx *ctx, uint32_t *out_dev, float *out_f32_dev)
{
    if (!ctx || !out_dev) return GUDNI_ERR_ARG;
    if (!ctx->frame_ready) return set_error(ctx, GUDNI_ERR_ARG, "gudni_frame_upload was not called");
    if (ctx->frame_n_work > kStaticMaxWork)
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_frame_render_static: the frame has " +
                                                 std::to_string(ctx->frame_n_work) +
                                                 " work units; the static scratch holds at most " +
                                                 std::to_string(kStaticMaxWork));
    if (ctx->spec.max_thresholds > (1 << kStaticSlowLog2))
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_frame_render_static: max_thresholds above " +
                                                 std::to_string(1 << kStaticSlowLog2) +
                                                 " (the static retry ring's capacity)");
    hipError_t err;
    uint8_t *base = static_scratch(ctx->device, err);
    if (!base)
        return set_error(ctx, GUDNI_ERR_HIP, std::string("gudni_frame_render_static: no address for the "
                                                         "static scratch on this device: ") +
                                                 hipGetErrorString(err));
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    dev::Frame f = base_frame(ctx);
    f.tiles = (const TileRec *)ctx->tiles.p;
    f.tile_job = (const int32_t *)ctx->tile_job.p;
    f.work = (const uint32_t *)ctx->work.p;
    f.n_work = ctx->frame_n_work;
    dev::SceneDesc sd = single_desc(ctx, ctx->background, ctx->frame_w, ctx->frame_h, out_dev, out_f32_dev);
    sd.sort_bits = ctx->frame_sort_bits;
    dev::SceneDesc *sdev = reinterpret_cast<dev::SceneDesc *>(base + kOffScene);
    hipLaunchKernelGGL(dev::store_scene, dim3(1), dim3(64), 0, ctx->stream, sdev, sd);
    CHECK_HIP(ctx, hipGetLastError());
    f.scenes = sdev;
    f.tile_scene = nullptr;
    f.scene_done = nullptr;
    ctx->timed_render = false;   // timing would need events and a cycle buffer
    const PassBuffers b{reinterpret_cast<uint32_t *>(base + kOffCtl), kStaticMaxWork,
                        reinterpret_cast<float4 *>(base + kOffFastThr),
                        reinterpret_cast<uint32_t *>(base + kOffFastHdr),
                        reinterpret_cast<int32_t *>(base + kOffFastSidx), kStaticFastBlocks,
                        reinterpret_cast<float4 *>(base + kOffSlowThr),
                        reinterpret_cast<uint32_t *>(base + kOffSlowHdr),
                        reinterpret_cast<int32_t *>(base + kOffSlowSidx), 1, 64};
    return launch_passes(ctx, f, b);
}

int gudni_frame_check(gudni_ctx *ctx)
{
    if (!ctx) return GUDNI_ERR_ARG;
    int rc = read_status(ctx);
    if (rc) return rc;
    if (ctx->timed_render && ctx->ev[0] && ctx->ev[1] && ctx->ev[2]) {
        float a = 0, b = 0;
        CHECK_HIP(ctx, hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]));
        CHECK_HIP(ctx, hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]));
        ctx->last_ms[0] = a;
        ctx->last_ms[1] = b;
        CHECK_HIP(ctx, hipMemcpy(ctx->last_cycles, ctx->cycles.p, sizeof(ctx->last_cycles),
                                 hipMemcpyDeviceToHost));
        ctx->timed_render = false;
    }
    return GUDNI_OK;
}

int gudni_scene_upload(gudni_ctx *ctx, const gudni_scene *sc)
{
    if (!ctx || !sc) return GUDNI_ERR_ARG;
    if (sc->n_shapes < 0 || sc->width <= 0 || sc->height <= 0 || (sc->n_shapes && (!sc->shapes || !sc->bbox)))
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_scene_upload: bad scene");
    if (std::max(sc->width, sc->height) > (1 << 15))
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_scene_upload: canvas side above 32768");
    ctx->scene_ready = false;
    int rc = gudni_upload_common(ctx, sc->geo, sc->geo_bytes, sc->substances, sc->n_substances,
                                 sc->pict_data, sc->pict_bytes, sc->pict_uses, sc->n_pict_uses,
                                 sc->random_field, sc->n_random);
    if (rc) return rc;
    rc = gudni_upload_gradients(ctx, sc->gradients, sc->n_gradients);
    if (rc) return rc;
    std::vector<float> bounds(4 * (size_t)sc->n_shapes + 4);
    rc = validate_job(ctx, ctx_view(ctx), (const ShapeRec *)sc->shapes, sc->n_shapes, nullptr, 0,
                      sc->width, sc->height, bounds);
    if (rc) return rc;
    CHECK_HIP(ctx, upload(ctx, ctx->scene_shapes, (const uint8_t *)sc->shapes, 16u * (size_t)sc->n_shapes));
    CHECK_HIP(ctx, upload(ctx, ctx->scene_bbox, (const uint8_t *)sc->bbox, 16u * (size_t)sc->n_shapes));
    CHECK_HIP(ctx, upload(ctx, ctx->scene_bounds, bounds.data(), 16u * (size_t)sc->n_shapes));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->scene_n_shapes = sc->n_shapes;
    ctx->scene_geo_bytes = sc->geo_bytes;
    ctx->scene_w = sc->width;
    ctx->scene_h = sc->height;
    std::memcpy(ctx->scene_bg, sc->background, sizeof(ctx->scene_bg));
    ctx->scene_ready = true;
    return GUDNI_OK;
}

int gudni_scene_bin(gudni_ctx *ctx)
{
    if (!ctx) return GUDNI_ERR_ARG;
    if (!ctx->scene_ready) return set_error(ctx, GUDNI_ERR_ARG, "gudni_scene_upload was not called");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    int rc = ensure_scratch(ctx);
    if (rc) return rc;
    const BinIn in{(const ShapeRec *)ctx->scene_shapes.p, (const float4 *)ctx->scene_bbox.p,
                   (const float4 *)ctx->scene_bounds.p, ctx->scene_n_shapes, ctx->scene_w, ctx->scene_h};
    const BinOut out{&ctx->tiles, &ctx->tile_job, &ctx->shapes, &ctx->xrange, &ctx->work};
    BinCounts cnt;
    rc = bin_scene(ctx, in, out, cnt);
    if (rc) return rc;
    const uint32_t status[4] = {0u, (uint32_t)cnt.n_tiles, (uint32_t)cnt.n_refs, (uint32_t)cnt.n_work};
    rc = ensure_control(ctx, (int)status[3]);
    if (rc) return rc;
    ctx->frame_w = ctx->scene_w;
    ctx->frame_h = ctx->scene_h;
    ctx->frame_n_tiles = (int32_t)status[1];
    ctx->scene_n_refs = (int32_t)status[2];
    ctx->frame_n_work = (int32_t)status[3];
    ctx->frame_index = 0;
    ctx->frame_sort_bits = ctx->spec.sort_radix_bits;
    std::memcpy(ctx->background, ctx->scene_bg, sizeof(ctx->background));
    ctx->frame_ready = true;
    return GUDNI_OK;
}

int gudni_scene_counts(gudni_ctx *ctx, int32_t counts[3])
{
    if (!ctx || !counts) return GUDNI_ERR_ARG;
    if (!ctx->frame_ready || !ctx->scene_ready)
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_scene_bin was not called");
    counts[0] = ctx->frame_n_tiles;
    counts[1] = ctx->scene_n_refs;
    counts[2] = ctx->frame_n_work;
    return GUDNI_OK;
}

int gudni_scene_read(gudni_ctx *ctx, void *tiles, int32_t *tile_job, void *refs, float *ref_bounds,
                     uint32_t *work)
{
    int32_t c[3];
    const int rc = gudni_scene_counts(ctx, c);
    if (rc) return rc;
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (tiles) CHECK_HIP(ctx, hipMemcpy(tiles, ctx->tiles.p, 32u * (size_t)c[0], hipMemcpyDeviceToHost));
    if (tile_job) CHECK_HIP(ctx, hipMemcpy(tile_job, ctx->tile_job.p, 4u * (size_t)c[0], hipMemcpyDeviceToHost));
    if (refs) CHECK_HIP(ctx, hipMemcpy(refs, ctx->shapes.p, 16u * (size_t)c[1], hipMemcpyDeviceToHost));
    if (ref_bounds) CHECK_HIP(ctx, hipMemcpy(ref_bounds, ctx->xrange.p, 16u * (size_t)c[1], hipMemcpyDeviceToHost));
    if (work) CHECK_HIP(ctx, hipMemcpy(work, ctx->work.p, 4u * (size_t)c[2], hipMemcpyDeviceToHost));
    return GUDNI_OK;
}

float gudni_last_bin_ms(const gudni_ctx *ctx) { return ctx ? ctx->last_bin_ms : 0.0f; }

int gudni_raster_scene(gudni_ctx *ctx, const gudni_scene *sc, uint32_t *host_out)
{
    if (!ctx || !sc || !host_out) return GUDNI_ERR_ARG;
    int rc = gudni_scene_upload(ctx, sc);
    if (rc) return rc;
    rc = gudni_scene_bin(ctx);
    if (rc) return rc;
    const size_t npix = (size_t)sc->width * (size_t)sc->height;
    CHECK_HIP(ctx, ctx->out.ensure(4 * npix));
    rc = gudni_frame_render(ctx, (uint32_t *)ctx->out.p, nullptr);
    if (rc) return rc;
    rc = gudni_frame_check(ctx);
    if (rc) return rc;
    CHECK_HIP(ctx, hipMemcpy(host_out, ctx->out.p, 4 * npix, hipMemcpyDeviceToHost));
    return GUDNI_OK;
}

int gudni_set_output_format(gudni_ctx *ctx, int32_t format)
{
    if (!ctx) return GUDNI_ERR_ARG;
    if (format != GUDNI_FORMAT_BGRA8 && format != GUDNI_FORMAT_RGBA8)
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_set_output_format: GUDNI_FORMAT_BGRA8 or GUDNI_FORMAT_RGBA8");
    ctx->out_format = format;
    return GUDNI_OK;
}

int gudni_set_blocks_per_cu(gudni_ctx *ctx, int32_t blocks_per_cu)
{
    if (!ctx) return GUDNI_ERR_ARG;
    if (blocks_per_cu < 0 || blocks_per_cu > 8)
        return set_error(ctx, GUDNI_ERR_ARG, "blocks per CU must be 0 (as many as fit) or 1..8");
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->blocks_per_cu = blocks_per_cu;
    ctx->fast_blocks = 0;   // resized on the next launch
    return ensure_scratch(ctx);
}

int gudni_enable_timing(gudni_ctx *ctx, int enable)
{
    if (!ctx) return GUDNI_ERR_ARG;
    ctx->timing = enable != 0;
    return GUDNI_OK;
}

int gudni_last_timings(gudni_ctx *ctx, float *ms, int32_t n)
{
    if (!ctx || !ms) return GUDNI_ERR_ARG;
    for (int i = 0; i < n && i < 2; ++i) ms[i] = ctx->last_ms[i];
    return GUDNI_OK;
}

int gudni_last_stage_cycles(gudni_ctx *ctx, uint64_t *cycles, int32_t n)
{
    if (!ctx || !cycles) return GUDNI_ERR_ARG;
    for (int i = 0; i < n && i < 8; ++i) cycles[i] = ctx->last_cycles[i];
    return GUDNI_OK;
}


// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------
struct gudni_batch {
    int32_t n_scenes = 0;
    int device = 0;
    DevBuf geo, subs, pict, uses, shapes, bounds, tiles, tile_job, tile_scene, work, scenes, done,
        frames, random, grads;
    DevBuf work_all;   // the whole-batch launch's work list, longest units first
    std::vector<size_t> frame_off;
    std::vector<int32_t> width, height;
    std::vector<int32_t> work_begin, work_count;
    int32_t n_work = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t captured_gen = 0;     // ctx->scratch_gen the graph was captured with
    int32_t captured_format = 0;
    hipStream_t capture_stream = nullptr;   // capture needs a stream other than the null stream
};


int gudni_batch_upload(gudni_ctx *ctx, int32_t n_scenes, const gudni_batch_scene *scenes,
                       const float *random_field, int32_t n_random, gudni_batch **out)
{
    if (!ctx || !scenes || !out || n_scenes <= 0) return GUDNI_ERR_ARG;
    if (!random_field || n_random != ctx->spec.random_field_size)
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_batch_upload: random field must hold "
                                             "random_field_size floats");
    *out = nullptr;
    std::vector<uint8_t> geo, pict, uses;
    std::vector<float4> subs;
    std::vector<ShapeRec> refs;
    std::vector<float> bounds;
    std::vector<TileRec> tiles;
    std::vector<int32_t> tile_job, tile_scene, wbegin, wcount;
    std::vector<uint32_t> work;
    std::vector<size_t> geo_off, subs_off, pict_off, uses_off, refs_off, frame_off, grads_off;
    std::vector<gudni_gradient> grads;
    size_t frame_words = 0;
    for (int32_t s = 0; s < n_scenes; ++s) {
        const gudni_batch_scene &sc = scenes[s];
        if (!sc.jobs || sc.width <= 0 || sc.height <= 0 || sc.n_substances < 0 ||
            sc.n_pict_uses < 0 || (sc.geo_bytes && !sc.geo) || (sc.n_substances && !sc.substances) ||
            (sc.pict_bytes && !sc.pict_data) || (sc.n_pict_uses && !sc.pict_uses))
            return set_error(ctx, GUDNI_ERR_ARG, "gudni_batch_upload: bad scene " + std::to_string(s));
        int rc = validate_uses(ctx, sc.pict_uses, sc.n_pict_uses, sc.pict_bytes);
        if (rc) return rc;
        rc = validate_gradients(ctx, sc.gradients, sc.n_gradients);
        if (rc) return rc;
        gudni_jobs_view v;
        gudni_jobs_view_get(sc.jobs, &v);
        std::vector<float4> sub_host((size_t)sc.n_substances);
        if (sc.n_substances) std::memcpy(sub_host.data(), sc.substances, 16u * (size_t)sc.n_substances);
        const CommonView cv{(const uint8_t *)sc.geo, sc.geo_bytes, sub_host.data(), sc.n_substances,
                            sc.n_pict_uses, sc.n_gradients};
        std::vector<TileRec> t(( const TileRec *)v.tiles, (const TileRec *)v.tiles + v.n_tiles);
        std::vector<int32_t> tj((size_t)v.n_tiles);
        for (int j = 0; j < v.n_jobs; ++j)
            for (int i = v.job_tile_offset[j]; i < v.job_tile_offset[j + 1]; ++i) {
                t[i].shape_start += v.job_shape_offset[j];
                tj[i] = v.n_jobs - 1 - j;
            }
        std::vector<float> xr(4 * (size_t)v.n_refs + 4);
        rc = validate_job(ctx, cv, (const ShapeRec *)v.shape_refs, v.n_refs, t.data(), v.n_tiles,
                          sc.width, sc.height, xr);
        if (rc) return set_error(ctx, rc, "scene " + std::to_string(s) + ": " + ctx->error);
        geo_off.push_back(geo.size());
        geo.insert(geo.end(), (const uint8_t *)sc.geo, (const uint8_t *)sc.geo + sc.geo_bytes);
        geo.resize(align_up(geo.size(), 16));
        subs_off.push_back(subs.size());
        subs.insert(subs.end(), sub_host.begin(), sub_host.end());
        grads_off.push_back(grads.size());
        if (sc.n_gradients)
            grads.insert(grads.end(), (const gudni_gradient *)sc.gradients,
                         (const gudni_gradient *)sc.gradients + sc.n_gradients);
        pict_off.push_back(pict.size());
        if (sc.pict_bytes)
            pict.insert(pict.end(), (const uint8_t *)sc.pict_data, (const uint8_t *)sc.pict_data + sc.pict_bytes);
        pict.resize(align_up(pict.size(), 16));
        uses_off.push_back(uses.size() / 24);
        if (sc.n_pict_uses)
            uses.insert(uses.end(), (const uint8_t *)sc.pict_uses,
                        (const uint8_t *)sc.pict_uses + 24 * (size_t)sc.n_pict_uses);
        refs_off.push_back(refs.size());
        refs.insert(refs.end(), (const ShapeRec *)v.shape_refs, (const ShapeRec *)v.shape_refs + v.n_refs);
        bounds.insert(bounds.end(), xr.begin(), xr.begin() + 4 * (size_t)v.n_refs);
        const int32_t tile_base = (int32_t)tiles.size();
        std::vector<uint32_t> w;
        build_work(ctx->spec, t.data(), v.n_tiles, sc.width, sc.height, w);
        wbegin.push_back((int32_t)work.size());
        wcount.push_back((int32_t)w.size());
        for (uint32_t u : w) work.push_back(u + ((uint32_t)tile_base << 2));
        tiles.insert(tiles.end(), t.begin(), t.end());
        tile_job.insert(tile_job.end(), tj.begin(), tj.end());
        tile_scene.insert(tile_scene.end(), (size_t)v.n_tiles, s);
        frame_off.push_back(frame_words);
        frame_words += (size_t)sc.width * sc.height;
    }
    if (tiles.size() >= (1u << 30))
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_batch_upload: too many tiles");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    int rc = ensure_scratch(ctx);
    if (rc) return rc;
    rc = ensure_control(ctx, (int)work.size());
    if (rc) return rc;
    gudni_batch *b = new gudni_batch();
    b->n_scenes = n_scenes;
    b->device = ctx->device;
    b->n_work = (int32_t)work.size();
    b->frame_off = frame_off;
    b->work_begin = wbegin;
    b->work_count = wcount;
    for (int32_t s = 0; s < n_scenes; ++s) {
        b->width.push_back(scenes[s].width);
        b->height.push_back(scenes[s].height);
    }
    auto up = [&](DevBuf &d, const void *src, size_t bytes) -> hipError_t {
        hipError_t e = d.ensure(bytes);
        if (e != hipSuccess) return e;
        return bytes ? hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = up(b->geo, geo.data(), geo.size());
    if (e == hipSuccess) e = up(b->subs, subs.data(), 16 * subs.size());
    if (e == hipSuccess) e = up(b->pict, pict.data(), pict.size());
    if (e == hipSuccess) e = up(b->uses, uses.data(), uses.size());
    if (e == hipSuccess) e = up(b->shapes, refs.data(), 16 * refs.size());
    if (e == hipSuccess) e = up(b->bounds, bounds.data(), 4 * bounds.size());
    if (e == hipSuccess) e = up(b->tiles, tiles.data(), 32 * tiles.size());
    if (e == hipSuccess) e = up(b->tile_job, tile_job.data(), 4 * tile_job.size());
    if (e == hipSuccess) e = up(b->tile_scene, tile_scene.data(), 4 * tile_scene.size());
    if (e == hipSuccess) e = up(b->work, work.data(), 4 * work.size());
    if (e == hipSuccess) {
        std::vector<uint32_t> all(work.size());
        order_longest_first(ctx->spec.threads_per_tile, tiles.data(), tile_scene.data(), work.data(),
                            (int)work.size(), dev::kParts, all.data());
        e = up(b->work_all, all.data(), 4 * all.size());
    }
    if (e == hipSuccess) e = up(b->random, random_field, 4 * (size_t)n_random);
    if (e == hipSuccess) e = up(b->grads, grads.data(), sizeof(gudni_gradient) * grads.size());
    if (e == hipSuccess) e = b->frames.ensure(4 * frame_words);
    if (e == hipSuccess) e = hipMemset(b->frames.p, 0, 4 * frame_words);
    if (e == hipSuccess) e = b->done.ensure(8 * ((size_t)n_scenes + 1));
    if (e == hipSuccess) e = hipMemset(b->done.p, 0, 8 * ((size_t)n_scenes + 1));
    std::vector<dev::SceneDesc> descs((size_t)n_scenes);
    for (int32_t s = 0; s < n_scenes && e == hipSuccess; ++s) {
        dev::SceneDesc &d = descs[s];
        std::memset(&d, 0, sizeof(d));
        d.geo = (const uint8_t *)b->geo.p + geo_off[s];
        d.shapes = (const ShapeRec *)b->shapes.p + refs_off[s];
        d.bounds = (const float4 *)b->bounds.p + refs_off[s];
        d.substances = (const float4 *)b->subs.p + subs_off[s];
        d.pict = (const uint8_t *)b->pict.p + pict_off[s];
        d.uses = (const PictUse *)b->uses.p + uses_off[s];
        d.out = (uint32_t *)b->frames.p + frame_off[s];
        d.out_f32 = nullptr;
        d.grads = (const float4 *)((const gudni_gradient *)b->grads.p + grads_off[s]);
        d.filter = ctx->spec.texture_filter;
        d.sort_bits = scenes[s].jobs->j.sort_radix_bits;
        for (int i = 0; i < 4; ++i) d.bg[i] = scenes[s].background[i];
        d.width = scenes[s].width;
        d.height = scenes[s].height;
    }
    if (e == hipSuccess) e = up(b->scenes, descs.data(), sizeof(dev::SceneDesc) * descs.size());
    if (e != hipSuccess) {
        ctx->error = std::string("gudni_batch_upload: ") + hipGetErrorString(e);
        gudni_batch_free(b);
        return GUDNI_ERR_HIP;
    }
    *out = b;
    return GUDNI_OK;
}


int gudni_batch_upload_outlines(gudni_ctx *ctx, int32_t n_scenes, const gudni_outlines *outlines,
                                const gudni_batch_scene *scenes, const int32_t *tile_sizes,
                                const int32_t *radix_bits, const float *random_field, int32_t n_random,
                                gudni_batch **out)
{
    if (!ctx || !outlines || !scenes || !out || n_scenes <= 0) return GUDNI_ERR_ARG;
    if (!random_field || n_random != ctx->spec.random_field_size)
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_batch_upload_outlines: random field must hold "
                                             "random_field_size floats");
    *out = nullptr;
    std::vector<uint8_t> pict, uses;
    std::vector<float4> subs;
    std::vector<gudni_gradient> grads;
    std::vector<size_t> subs_off, pict_off, uses_off, grads_off, frame_off;
    size_t frame_words = 0;
    for (int32_t s = 0; s < n_scenes; ++s) {
        const gudni_batch_scene &sc = scenes[s];
        const std::string where = "gudni_batch_upload_outlines: scene " + std::to_string(s);
        if (sc.jobs || sc.geo || sc.geo_bytes || sc.width <= 0 || sc.height <= 0 || sc.n_substances < 0 ||
            sc.n_pict_uses < 0 || (sc.n_substances && !sc.substances) || (sc.pict_bytes && !sc.pict_data) ||
            (sc.n_pict_uses && !sc.pict_uses) || std::max(sc.width, sc.height) > (1 << 15))
            return set_error(ctx, GUDNI_ERR_ARG, where + ": bad scene (geometry and jobs must be empty)");
        const int32_t ts = tile_sizes ? tile_sizes[s] : ctx->spec.max_tile_size;
        const int32_t rb = radix_bits ? radix_bits[s] : ctx->spec.sort_radix_bits;
        if (ts < 8 || ts > 256 || (ts & (ts - 1)) || rb < 0 || rb > dev::kRadixMaxBits)
            return set_error(ctx, GUDNI_ERR_ARG, where + ": tile size must be a power of two in 8..256 "
                                                         "and the sort radix 0..5");
        const gudni_outlines &o = outlines[s];
        if (o.n_shapes < 0 || (o.n_shapes && (!o.tags || !o.shape_outline_offsets || !o.outline_pair_offsets)))
            return set_error(ctx, GUDNI_ERR_ARG, where + ": bad outlines");
        int32_t n_outlines = 0, n_pairs = 0;
        int rc = check_outline_offsets(ctx, &o, n_outlines, n_pairs);
        if (rc) return set_error(ctx, rc, where + ": " + ctx->error);
        rc = validate_uses(ctx, sc.pict_uses, sc.n_pict_uses, sc.pict_bytes);
        if (rc) return rc;
        rc = validate_gradients(ctx, sc.gradients, sc.n_gradients);
        if (rc) return rc;
        subs_off.push_back(subs.size());
        if (sc.n_substances)
            subs.insert(subs.end(), (const float4 *)sc.substances, (const float4 *)sc.substances + sc.n_substances);
        grads_off.push_back(grads.size());
        if (sc.n_gradients)
            grads.insert(grads.end(), (const gudni_gradient *)sc.gradients,
                         (const gudni_gradient *)sc.gradients + sc.n_gradients);
        pict_off.push_back(pict.size());
        if (sc.pict_bytes)
            pict.insert(pict.end(), (const uint8_t *)sc.pict_data, (const uint8_t *)sc.pict_data + sc.pict_bytes);
        pict.resize(align_up(pict.size(), 16));
        uses_off.push_back(uses.size() / 24);
        if (sc.n_pict_uses)
            uses.insert(uses.end(), (const uint8_t *)sc.pict_uses,
                        (const uint8_t *)sc.pict_uses + 24 * (size_t)sc.n_pict_uses);
        frame_off.push_back(frame_words);
        frame_words += (size_t)sc.width * sc.height;
    }
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    int rc = ensure_scratch(ctx);
    if (rc) return rc;
    const hipStream_t st = ctx->stream;
    gudni_batch *b = new gudni_batch();
    b->n_scenes = n_scenes;
    b->device = ctx->device;
    b->frame_off = frame_off;
    for (int32_t s = 0; s < n_scenes; ++s) {
        b->width.push_back(scenes[s].width);
        b->height.push_back(scenes[s].height);
    }
    auto fail = [&](int code, const std::string &msg) {
        if (!msg.empty()) ctx->error = msg;
        gudni_batch_free(b);
        return code;
    };
    auto hip_fail = [&](hipError_t e) {
        return fail(GUDNI_ERR_HIP, std::string("gudni_batch_upload_outlines: ") + hipGetErrorString(e));
    };
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = upload(ctx, b->subs, subs.data(), 16 * subs.size());
    if (e == hipSuccess) e = upload(ctx, b->pict, pict.data(), pict.size());
    if (e == hipSuccess) e = upload(ctx, b->uses, uses.data(), uses.size());
    if (e == hipSuccess) e = upload(ctx, b->grads, grads.data(), sizeof(gudni_gradient) * grads.size());
    if (e == hipSuccess) e = upload(ctx, b->random, random_field, 4 * (size_t)n_random);
    if (e == hipSuccess) e = b->frames.ensure(4 * frame_words);
    if (e == hipSuccess) e = hipMemsetAsync(b->frames.p, 0, 4 * frame_words, st);
    if (e == hipSuccess) e = b->done.ensure(8 * ((size_t)n_scenes + 1));
    if (e == hipSuccess) e = hipMemsetAsync(b->done.p, 0, 8 * ((size_t)n_scenes + 1), st);
    if (e != hipSuccess) return hip_fail(e);
    // every scene serialised and binned on the device, appended to the batch
    size_t geo_used = 0, refs_used = 0, bounds_used = 0, tiles_used = 0, tj_used = 0, ts_used = 0, work_used = 0;
    std::vector<size_t> geo_off, refs_off;
    ctx->scene_ready = false;
    ctx->frame_ready = false;
    for (int32_t s = 0; s < n_scenes; ++s) {
        const gudni_outlines &o = outlines[s];
        const gudni_batch_scene &sc = scenes[s];
        const int32_t n = o.n_shapes;
        const int32_t n_outlines = n ? o.shape_outline_offsets[n] : 0;
        const int32_t n_pairs = n_outlines ? o.outline_pair_offsets[n_outlines] : 0;
        if (e == hipSuccess) e = upload(ctx, ctx->ser_tags, o.tags, 8u * (size_t)n);
        if (e == hipSuccess) e = upload(ctx, ctx->ser_soff, o.shape_outline_offsets, 4u * ((size_t)n + 1));
        if (e == hipSuccess) e = upload(ctx, ctx->ser_ooff, o.outline_pair_offsets, 4u * ((size_t)n_outlines + 1));
        if (e == hipSuccess) e = upload(ctx, ctx->ser_pairs, o.pairs, 16u * (size_t)n_pairs);
        if (e != hipSuccess) return hip_fail(e);
        SerIn in;
        in.tags = (const uint64_t *)ctx->ser_tags.p;
        in.soff = (const int32_t *)ctx->ser_soff.p;
        in.ooff = (const int32_t *)ctx->ser_ooff.p;
        in.pairs = (const float *)ctx->ser_pairs.p;
        in.n_shapes = n;
        in.width = sc.width;
        in.height = sc.height;
        in.subs = (const float4 *)b->subs.p + subs_off[s];
        in.n_subs = sc.n_substances;
        in.n_uses = sc.n_pict_uses;
        in.n_grads = sc.n_gradients;
        const SerOut so{&ctx->geo, &ctx->scene_shapes, &ctx->scene_bbox, &ctx->scene_bounds, nullptr};
        SerCounts sc_cnt;
        rc = serialize_outlines(ctx, in, so, sc_cnt);
        if (rc) return fail(rc, "scene " + std::to_string(s) + ": " + ctx->error);
        BinIn bi{(const ShapeRec *)ctx->scene_shapes.p, (const float4 *)ctx->scene_bbox.p,
                 (const float4 *)ctx->scene_bounds.p, (int32_t)sc_cnt.n_kept, sc.width, sc.height};
        bi.max_tile_size = tile_sizes ? tile_sizes[s] : 0;
        const BinOut bo{&ctx->tiles, &ctx->tile_job, &ctx->shapes, &ctx->xrange, &ctx->work};
        BinCounts bc;
        rc = bin_scene(ctx, bi, bo, bc);
        if (rc) return fail(rc, "scene " + std::to_string(s) + ": " + ctx->error);
        size_t off = 0;
        if (e == hipSuccess) e = append_dev(b->geo, geo_used, ctx->geo.p, (size_t)sc_cnt.geo_bytes, 16, st, off);
        geo_off.push_back(off);
        if (e == hipSuccess) e = append_dev(b->shapes, refs_used, ctx->shapes.p, 16u * (size_t)bc.n_refs, 16, st, off);
        refs_off.push_back(off / 16);
        if (e == hipSuccess)
            e = append_dev(b->bounds, bounds_used, ctx->xrange.p, 16u * (size_t)bc.n_refs, 16, st, off);
        const size_t tile_base = tiles_used / 32;
        if (e == hipSuccess) e = append_dev(b->tiles, tiles_used, ctx->tiles.p, 32u * (size_t)bc.n_tiles, 32, st, off);
        if (e == hipSuccess) e = append_dev(b->tile_job, tj_used, ctx->tile_job.p, 4u * (size_t)bc.n_tiles, 4, st, off);
        if (e == hipSuccess && bc.n_tiles) {
            e = append_dev(b->tile_scene, ts_used, ctx->tile_job.p, 4u * (size_t)bc.n_tiles, 4, st, off);
            if (e == hipSuccess)
                hipLaunchKernelGGL(fill_i32, dim3(std::min(1024, (bc.n_tiles + 255) / 256)), dim3(256), 0, st,
                                   (int32_t *)((uint8_t *)b->tile_scene.p + off), bc.n_tiles, (int32_t)s);
        }
        b->work_begin.push_back((int32_t)(work_used / 4));
        b->work_count.push_back(bc.n_work);
        if (e == hipSuccess && bc.n_work) {
            e = append_dev(b->work, work_used, ctx->work.p, 4u * (size_t)bc.n_work, 4, st, off);
            if (e == hipSuccess)
                hipLaunchKernelGGL(rebase_work, dim3(std::min(1024, (bc.n_work + 255) / 256)), dim3(256), 0, st,
                                   (uint32_t *)((uint8_t *)b->work.p + off), bc.n_work, (uint32_t)tile_base);
        }
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e);
        // the next scene's serialisation reuses the scratch: wait for the copies
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e);
    }
    if (tiles_used / 32 >= (1u << 30)) return fail(GUDNI_ERR_ARG, "gudni_batch_upload_outlines: too many tiles");
    for (DevBuf *d : {&b->geo, &b->shapes, &b->bounds, &b->tiles, &b->tile_job, &b->tile_scene, &b->work})
        if (e == hipSuccess && !d->p) e = d->ensure(16);
    if (e != hipSuccess) return hip_fail(e);
    b->n_work = (int32_t)(work_used / 4);
    {
        std::vector<TileRec> th(tiles_used / 32);
        std::vector<int32_t> ts(tiles_used / 32);
        std::vector<uint32_t> wh((size_t)b->n_work), all((size_t)b->n_work);
        if (!th.empty()) e = hipMemcpy(th.data(), b->tiles.p, tiles_used, hipMemcpyDeviceToHost);
        if (e == hipSuccess && !ts.empty())
            e = hipMemcpy(ts.data(), b->tile_scene.p, 4 * ts.size(), hipMemcpyDeviceToHost);
        if (e == hipSuccess && !wh.empty()) e = hipMemcpy(wh.data(), b->work.p, work_used, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e);
        order_longest_first(ctx->spec.threads_per_tile, th.data(), ts.data(), wh.data(), b->n_work, dev::kParts,
                            all.data());
        e = upload(ctx, b->work_all, all.data(), 4 * all.size());
        if (e != hipSuccess) return hip_fail(e);
    }
    rc = ensure_control(ctx, b->n_work);
    if (rc) return fail(rc, "");
    std::vector<dev::SceneDesc> descs((size_t)n_scenes);
    for (int32_t s = 0; s < n_scenes; ++s) {
        dev::SceneDesc &d = descs[s];
        std::memset(&d, 0, sizeof(d));
        d.geo = (const uint8_t *)b->geo.p + geo_off[s];
        d.shapes = (const ShapeRec *)b->shapes.p + refs_off[s];
        d.bounds = (const float4 *)b->bounds.p + refs_off[s];
        d.substances = (const float4 *)b->subs.p + subs_off[s];
        d.pict = (const uint8_t *)b->pict.p + pict_off[s];
        d.uses = (const PictUse *)b->uses.p + uses_off[s];
        d.out = (uint32_t *)b->frames.p + frame_off[s];
        d.out_f32 = nullptr;
        d.grads = (const float4 *)((const gudni_gradient *)b->grads.p + grads_off[s]);
        d.filter = ctx->spec.texture_filter;
        d.sort_bits = radix_bits ? radix_bits[s] : ctx->spec.sort_radix_bits;
        for (int i = 0; i < 4; ++i) d.bg[i] = scenes[s].background[i];
        d.width = scenes[s].width;
        d.height = scenes[s].height;
    }
    e = upload(ctx, b->scenes, descs.data(), sizeof(dev::SceneDesc) * descs.size());
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e);
    *out = b;
    return GUDNI_OK;
}

static dev::Frame batch_frame(gudni_ctx *ctx, gudni_batch *b)
{
    dev::Frame f = base_frame(ctx);
    f.scenes = (const dev::SceneDesc *)b->scenes.p;
    f.tile_scene = (const int32_t *)b->tile_scene.p;
    f.scene_done = (unsigned long long *)b->done.p + 1;
    f.tiles = (const TileRec *)b->tiles.p;
    f.tile_job = (const int32_t *)b->tile_job.p;
    f.work = (const uint32_t *)b->work_all.p;
    f.n_work = b->n_work;
    f.random = (const float *)b->random.p;
    return f;
}

static int batch_launch(gudni_ctx *ctx, gudni_batch *b)
{
    const int threads = 256, blocks = (b->n_scenes + 1 + threads - 1) / threads;
    hipLaunchKernelGGL(dev::batch_begin, dim3(blocks), dim3(threads), 0, ctx->stream,
                       (unsigned long long *)b->done.p, b->n_scenes);
    CHECK_HIP(ctx, hipGetLastError());
    return launch_passes(ctx, batch_frame(ctx, b));
}

int gudni_batch_render(gudni_ctx *ctx, gudni_batch *b, int use_graph)
{
    if (!ctx || !b) return GUDNI_ERR_ARG;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    int rc = ensure_control(ctx, b->n_work);
    if (rc) return rc;
    ctx->timed_render = false;
    if (!use_graph) return batch_launch(ctx, b);
    // the graph holds the scratch pointers, grid size and pixel format it
    // was captured with: re-capture once any of them changed
    if (b->exec && (b->captured_gen != ctx->scratch_gen || b->captured_format != ctx->out_format)) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
    }
    if (!b->exec) {
        // The launches are recorded on the batch's own stream and the graph is
        // replayed on the context's stream, which may be the null stream.
        if (!b->capture_stream)
            CHECK_HIP(ctx, hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
        CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const hipStream_t user = ctx->stream;
        ctx->stream = b->capture_stream;
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            ctx->stream = user;
            CHECK_HIP(ctx, e);
        }
        rc = batch_launch(ctx, b);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(ctx->stream, &g);
        ctx->stream = user;
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        CHECK_HIP(ctx, e);
        b->graph = g;
        CHECK_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->captured_gen = ctx->scratch_gen;
        b->captured_format = ctx->out_format;
    }
    // the replay writes the status word of the context's control words
    ctx->status_word = (uint32_t *)ctx->control.p + dev::kCtlStatus;
    CHECK_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
    return GUDNI_OK;
}

int gudni_batch_render_scene(gudni_ctx *ctx, gudni_batch *b, int32_t scene)
{
    if (!ctx || !b || scene < 0 || scene >= b->n_scenes) return GUDNI_ERR_ARG;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    int rc = ensure_control(ctx, b->n_work);
    if (rc) return rc;
    ctx->timed_render = false;
    dev::Frame f = batch_frame(ctx, b);
    f.work = (const uint32_t *)b->work.p + b->work_begin[scene];
    f.n_work = b->work_count[scene];
    f.scene_done = nullptr;
    return launch_passes(ctx, f);
}

int gudni_batch_check(gudni_ctx *ctx, gudni_batch *b, float *latency_ms)
{
    if (!ctx || !b) return GUDNI_ERR_ARG;
    int rc = read_status(ctx);
    if (rc) return rc;
    if (latency_ms) {
        std::vector<unsigned long long> done((size_t)b->n_scenes + 1);
        CHECK_HIP(ctx, hipMemcpy(done.data(), b->done.p, 8 * done.size(), hipMemcpyDeviceToHost));
        int rate_khz = 0;
        CHECK_HIP(ctx, hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
        if (rate_khz <= 0) rate_khz = 100000;
        for (int32_t s = 0; s < b->n_scenes; ++s)
            latency_ms[s] = done[s + 1] > done[0] ? (float)((double)(done[s + 1] - done[0]) / rate_khz) : 0.0f;
    }
    return GUDNI_OK;
}

int gudni_batch_frame(gudni_ctx *ctx, gudni_batch *b, int32_t scene, uint32_t *host_out)
{
    if (!ctx || !b || !host_out || scene < 0 || scene >= b->n_scenes) return GUDNI_ERR_ARG;
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CHECK_HIP(ctx, hipMemcpy(host_out, (uint32_t *)b->frames.p + b->frame_off[scene],
                             4 * (size_t)b->width[scene] * b->height[scene], hipMemcpyDeviceToHost));
    return GUDNI_OK;
}

int gudni_order_work(const void *tiles, int32_t n_tiles, const int32_t *tile_scene, const uint32_t *work,
                     int32_t n_work, int32_t threads_per_tile, int32_t parts, uint32_t *out)
{
    if (n_tiles < 0 || n_work < 0 || parts < 1 || threads_per_tile < 1 || (n_work && (!tiles || !work || !out)))
        return GUDNI_ERR_ARG;
    const TileRec *t = (const TileRec *)tiles;
    for (int32_t i = 0; i < n_work; ++i) {
        if ((int64_t)(work[i] >> 2) >= n_tiles) return GUDNI_ERR_ARG;
        // the cost estimate shifts by the tile's depth: a tile is at most 2^15 wide
        const TileRec &r = t[work[i] >> 2];
        if (r.hdepth < 0 || r.hdepth > 15) return GUDNI_ERR_ARG;
    }
    order_longest_first(threads_per_tile, t, tile_scene, work, n_work, parts, out);
    return GUDNI_OK;
}

void gudni_batch_free(gudni_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    DevBuf *bufs[] = {&b->geo, &b->subs, &b->pict, &b->uses, &b->shapes, &b->bounds, &b->tiles,
                      &b->tile_job, &b->tile_scene, &b->work, &b->scenes, &b->done, &b->frames,
                      &b->random, &b->grads, &b->work_all};
    for (DevBuf *d : bufs) d->release();
    delete b;
}

} // extern "C"
