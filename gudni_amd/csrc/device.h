// device.h -- the per-column device stages shared by libgudni_hip's kernels:
// threshold generation, the per-column sort, the coverage / compositing sweep
// and the work-unit driver (see raster.hip for the launch structure).
//
// Numerics: compiled with -ffp-contract=off and correctly rounded FP32
// division, and every FP32 expression is evaluated in the order the
// reference writes it, so records, their order and the coverage sums are
// bit-identical to a strict (non-relaxed) evaluation of Kernels.cl.
// min/max are the OpenCL definitions: max(x,y) = x < y ? y : x,
// min(x,y) = y < x ? y : x.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "gudni_hip.h"
#include "internal.h"

namespace gudni {
namespace dev {

constexpr uint32_t kPositiveSlope = 0x80000000u;
constexpr uint32_t kPersist = 0x40000000u;
constexpr uint32_t kPersistAndSlope = 0xC0000000u;
constexpr uint32_t kPersistTop = 0xC0000000u;
constexpr uint32_t kPersistBottom = 0x40000000u;
constexpr uint32_t kUnpersist = 0xBFFFFFFFu;
constexpr uint32_t kShapeBitMask = 0x0FFFFFFFu;
// Shape labels are below 128 (maxShapes <= 127), so only the low seven bits
// of the shape-bit field carry them; the adjoint keeps a record's origin in
// the bits above.
constexpr uint32_t kLabelMask = 0x7Fu;
// Fill rule (extension): records of a nonzero-fill shape carry kNonzero, and
// kReversed when their strand runs against the outline (bits 29 and 28,
// which the reference leaves clear: its bit 29 "enable" is unused on this
// path and its shape bits never exceed 127).
constexpr uint32_t kNonzero = 0x20000000u;
constexpr uint32_t kReversed = 0x10000000u;
// Per-lane words of the shape scratch: [0, 128) shape indices of a general
// tile's bits, [128, 256) winding numbers of nonzero-fill shapes by label.
constexpr int kSidxWords = 256;
constexpr int kWindOffset = 128;
constexpr int kOriginShift = 7;
constexpr uint32_t kOriginMask = 0x3FFFFFu;   // 22 bits: bits 7..28
constexpr float kMinCrop = 0.2f;          // MINCROP (Kernels.cl:47)
constexpr float kFlatness = 0.25f;        // TAXICAB_FLATNESS (Kernels.cl:250)

constexpr uint64_t kTagTypeMask = 0xC000000000000000ull;
constexpr uint64_t kTagSolid = 0x8000000000000000ull;
constexpr uint64_t kTagCompoundMask = 0x3000000000000000ull;
constexpr uint64_t kTagContinue = 0x1000000000000000ull;
constexpr uint64_t kTagAdd = 0x2000000000000000ull;
constexpr uint64_t kTagIdMask = 0x07FFFFFFFFFFFFFFull;   // bit 59 is the fill rule
constexpr uint64_t kTagNonzero = GUDNI_TAG_NONZERO;
// Both type bits set: a gradient substance (an extension; the reference
// knows solids and pictures only).
constexpr uint64_t kTagGradient = 0xC000000000000000ull;

constexpr int kWave = 64;
constexpr int kFastCapLog2 = 7;           // 128-record rings in the fast pass
constexpr long kSectionGuard = 1L << 22;  // sweep-length bound per column: the
                                          // longest legitimate sweep is 256 rows
                                          // x (2*4096+1) sections = 2.1M
constexpr int kMaxBisect = 64;            // bifurcation bound per intersection

// status codes a lane can raise (bit flags in the launch's status word)
// Control words of a launch: per-part work counters of the fast pass, the
// retry pass's counter, the status word, the retry list (count, units).
constexpr int kParts = 8;                  // one work range per XCD
constexpr int kCtlFast = 0, kCtlSlow = kParts, kCtlStatus = kParts + 1, kCtlRetry = kParts + 2;
constexpr int kCtlWords = kCtlRetry + 1;   // words before the retry entries
constexpr uint32_t kStatusOverflow = 1u;   // live queue > maxThresholds
constexpr uint32_t kStatusGuard = 2u;      // section guard hit

struct Thr { float top, bottom, left, right; };

// The scene buffers are read-only for the whole launch.  Reading them through
// the constant address space says so, which lets wave-uniform reads (the
// shape records and strand headers every lane walks) go to the scalar unit.
#define GUDNI_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const GUDNI_CONST T *ro(const T *p)
{
    return (const GUDNI_CONST T *)(p);
}
__device__ __forceinline__ float4 ld4(const GUDNI_CONST uint8_t *p)
{
    const GUDNI_CONST float *q = reinterpret_cast<const GUDNI_CONST float *>(p);
    return make_float4(q[0], q[1], q[2], q[3]);
}
__device__ __forceinline__ float2 ld2(const GUDNI_CONST uint8_t *p)
{
    const GUDNI_CONST float *q = reinterpret_cast<const GUDNI_CONST float *>(p);
    return make_float2(q[0], q[1]);
}

__device__ __forceinline__ float cl_max(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ float cl_min(float x, float y) { return y < x ? y : x; }

__device__ __forceinline__ bool hdr_pos(uint32_t h) { return (h & kPositiveSlope) != 0u; }
__device__ __forceinline__ bool hdr_persist(uint32_t h) { return (h & kPersist) != 0u; }
__device__ __forceinline__ bool hdr_persist_top(uint32_t h) { return (h & kPersistAndSlope) == kPersistTop; }
__device__ __forceinline__ bool hdr_persist_bottom(uint32_t h) { return (h & kPersistAndSlope) == kPersistBottom; }

// thresholdInvertedSlope (Kernels.cl:240-246)
__device__ __forceinline__ float inv_slope(uint32_t h, const Thr &t)
{
    const float sgn = hdr_pos(h) ? 1.0f : -1.0f;
    return (t.top == t.bottom) ? FLT_MAX : (t.right - t.left) / (t.bottom - t.top) * sgn;
}
__device__ __forceinline__ float top_x(uint32_t h, const Thr &t) { return hdr_pos(h) ? t.left : t.right; }

// thresholdIntersectX (Kernels.cl:900-913)
__device__ __forceinline__ float intersect_x(uint32_t h, const Thr &t, float y)
{
    if (t.left == t.right) return t.left;
    const float sy = hdr_pos(h) ? t.top : t.bottom;
    return ((y - sy) * inv_slope(h, t)) + t.left;
}

// thresholdIsBelow (Kernels.cl:1079-1094): a sorts after b
__device__ __forceinline__ bool is_below(uint32_t ah, const Thr &a, uint32_t bh, const Thr &b)
{
    if (a.top > b.top) return true;
    if (a.top == b.top) {
        const float ax = top_x(ah, a), bx = top_x(bh, b);
        if (ax > bx) return true;
        if (ax == bx && inv_slope(ah, a) > inv_slope(bh, b)) return true;
    }
    return false;
}

// Launch-wide read-only inputs.
// One scene: its read-only buffers and its frame.
struct alignas(16) SceneDesc {
    const uint8_t *geo;           // strand heap
    const ShapeRec *shapes;       // shape references (tile shape_start indexes here)
    const float4 *bounds;         // per reference: strand extent (xmin, ymin, xmax, ymax)
    const float4 *substances;
    const uint8_t *pict;
    const PictUse *uses;
    uint32_t *out;                // BGRA words
    float4 *out_f32;              // optional pre-quantisation colours
    const float4 *grads;          // gradient records, 7 float4 each
    float bg[4];                  // background RGBA
    int32_t width, height;
    int32_t filter;               // picture sampling: 0 nearest, 1 bilinear
    int32_t sort_bits;            // radix digit width of the record sort (0: insertion only)
};

struct Frame {
    const SceneDesc *scenes;      // scene table (one entry for a single frame)
    const int32_t *tile_scene;    // scene of each tile (null: all scene 0)
    unsigned long long *scene_done; // wall clock when a scene's last unit ended
    const TileRec *tiles;         // tile records, shape_start absolute
    const int32_t *tile_job;      // job index per tile
    const uint32_t *work;         // work units: tile << 2 | quarter
    int32_t n_work;
    const float *random;
    int32_t compute_depth, max_shapes, max_thresholds, random_mask;
    float stochastic;
    int32_t coverage_mode;
    int32_t out_rgba;             // pixel words in R,G,B,A byte order, else the reference's B,G,R,A
    // scratch
    float4 *q_thr;                // [slots][64][cap]
    uint32_t *q_hdr;              // [slots][64][cap]
    int32_t *shape_idx;           // [slots][max_shapes][64]
    int32_t cap_log2;
    // control
    uint32_t *counter;            // work counters of this pass, one per part
    int32_t n_parts;              // parts the work list is cut into
    uint32_t *status;             // OR of lane status flags
    uint32_t *retry;              // [0] = count, [1..] = abandoned work units
    int32_t retry_capacity;
    // dump hooks (Level-1 test path), null in normal runs
    float4 *dump_pre_thr; uint32_t *dump_pre_hdr;
    float4 *dump_thr; uint32_t *dump_hdr;
    int32_t *dump_qlen; uint32_t *dump_bits; uint64_t *dump_stack; int32_t *dump_sidx;
    int32_t dump_cap, threads_per_tile;
    // stage cycle counters (null unless timing): [0] generate [1] sort
    // [2] render [3] work units, summed over waves
    unsigned long long *stage_cycles;
};

// A lane's threshold queue: a ring of its own in the wave slot's scratch.
// Logical element 0 is the front of the reference's ThresholdQueue
// (qSlice.sStart).  A ring's records are contiguous, so the few records at
// the front of a queue, which the sweep reads again and again, share cache
// lines (with the rings of a wave interleaved record by record, as the
// reference interleaves its threshold columns, the 4096^2 frame took 3.6%
// longer).
struct Queue {
    float4 *thr;      // the lane's ring (cap records)
    uint32_t *hdr;
    int head, len, mask, maxt;
    bool spilled;     // ring too small: abandon the unit
    bool overflow;    // live queue > maxThresholds (reference would corrupt)

    __device__ __forceinline__ int row(int i) const { return (head + i) & mask; }
    __device__ __forceinline__ Thr get(int i) const
    {
        const float4 v = thr[row(i)];
        return Thr{v.x, v.y, v.z, v.w};
    }
    __device__ __forceinline__ uint32_t get_h(int i) const { return hdr[row(i)]; }
    __device__ __forceinline__ void set(int i, uint32_t h, const Thr &t)
    {
        const int r = row(i);
        thr[r] = make_float4(t.top, t.bottom, t.left, t.right);
        hdr[r] = h;
    }
    // pushTopSlot (Kernels.cl:402-405)
    __device__ __forceinline__ bool push_slot()
    {
        if (len + 1 > maxt) { overflow = true; return false; }
        if (len + 1 > mask + 1) { spilled = true; return false; }
        head = (head - 1) & mask;
        ++len;
        return true;
    }
    __device__ __forceinline__ void pop() { head = (head + 1) & mask; --len; }
};

// Work claiming.  The list is cut into n_parts contiguous ranges; block b
// starts on range b % n_parts.  Workgroups are dispatched to the XCDs round
// robin, so with one range per XCD a wave first drains the range its own
// XCD owns -- neighbouring tiles share shapes, whose records and strands are
// then read through one L2 -- and afterwards helps with the others.  Every
// wave visits every range, so the grid drains completely.
struct Claimer {
    uint32_t *counters;
    int parts, n_work, part0, k, begin, end;
    __device__ void init(uint32_t *c, int n_parts, int n, int first)
    {
        counters = c;
        parts = n_parts;
        n_work = n;
        part0 = first % n_parts;
        k = 0;
        range();
    }
    __device__ void range()
    {
        const int part = (part0 + k) % parts;
        begin = (int)(((long long)n_work * part) / parts);
        end = (int)(((long long)n_work * (part + 1)) / parts);
    }
    // next work index (wave-uniform), or -1 once every range is drained
    __device__ int next(int lane)
    {
        while (k < parts) {
            uint32_t w = 0;
            if (lane == 0) w = atomicAdd(counters + (part0 + k) % parts, 1u);
            w = __builtin_amdgcn_readfirstlane(w);
            if ((int)w < end - begin) return begin + (int)w;
            if (++k < parts) range();
        }
        return -1;
    }
};

struct Column {
    int dx, dy;                   // threadDelta
    int int_height;
    float float_height;
    uint32_t unique;              // threadUnique (random field seed)
    int shape_start, num_shapes;
};

struct Trav { float2 right, left, ctl; float xpos; int index; };

__device__ __forceinline__ float2 mid_point(float2 a, float2 b)
{   // (1-T)*a + T*b with T = 0.5f (Kernels.cl:864-866)
    return make_float2(0.5f * a.x + 0.5f * b.x, 0.5f * a.y + 0.5f * b.y);
}

// bifurcateCurve + intersectCurve (Kernels.cl:1226-1258) + yIntercept (:818-820)
__device__ float intersect_curve(Trav t)
{
    if (!(t.left.x == t.ctl.x && t.left.y == t.ctl.y)) {
        // at most kMaxBisect halvings: a curve inside a 2^30-pixel canvas is
        // flat after ~35; beyond that FP32 can reach a fixed point where an
        // unbounded loop would never end
        for (int it = 0; it < kMaxBisect; ++it) {
            const float2 lm = mid_point(t.left, t.ctl);
            const float2 rm = mid_point(t.ctl, t.right);
            const float2 on = mid_point(lm, rm);
            const float flat = fabsf(t.ctl.x - on.x) + fabsf(t.ctl.y - on.y);
            if (!(flat > kFlatness)) break;
            if (t.xpos < on.x) { t.ctl = lm; t.right = on; }
            else { t.left = on; t.ctl = rm; }
        }
    }
    if (t.left.x == t.right.x) return cl_min(t.left.y, t.right.y);
    return (((t.right.y - t.left.y) / (t.right.x - t.left.x)) * (t.xpos - t.left.x)) + t.left.y;
}

// addThreshold + trimThresholdTop (Kernels.cl:982-1005, 1175-1220)
__device__ __forceinline__ void add_threshold(Queue &q, const Column &c, uint32_t h, Thr t,
                                              bool &added, bool &enclosed)
{
    const bool keep = hdr_persist(h) || ((t.bottom - t.top) >= kMinCrop);
    if (!keep) return;
    enclosed = enclosed || ((t.top <= 0.0f) && hdr_persist_top(h)) ||
               ((t.bottom <= 0.0f) && hdr_persist_bottom(h));
    if (t.top < c.float_height && t.bottom > 0.0f && t.left < 1.0f) {
        if (t.top <= 0.0f) {
            const float sx = intersect_x(h, t, 0.0f);
            if (hdr_pos(h)) { t = Thr{0.0f, t.bottom, sx, t.right}; h &= kUnpersist; }
            else { t = Thr{0.0f, t.bottom, t.left, sx}; }
        }
        if (t.right <= 0.0f) {
            enclosed = true;
        } else {
            added = true;
            if (q.push_slot()) q.set(0, h, t);
        }
    }
}

// addLineSegment (Kernels.cl:1131-1173)
__device__ __forceinline__ void add_line(Queue &q, const Column &c, float2 l, float2 r,
                                         uint32_t bit, bool &added, bool &enclosed)
{
    const Thr t{cl_min(l.y, r.y), cl_max(l.y, r.y), l.x, r.x};
    const bool pos = l.y <= r.y;
    const bool persistent = (l.x != r.x) && (l.x == 0.0f);
    const uint32_t h = (pos ? kPositiveSlope : 0u) | (persistent ? kPersist : 0u) | bit;
    add_threshold(q, c, h, t, added, enclosed);
}

// spawnThresholds (Kernels.cl:1264-1333)
__device__ void spawn(Queue &q, const Column &c, uint32_t bit, const Trav &l, const Trav &r,
                      bool &added, bool &enclosed)
{
    const float yl = (l.left.x >= 0.0f) ? l.left.y : intersect_curve(l);
    bool lw = false, rw = false;
    if ((l.right.x < 1.0f) && (l.right.x > 0.0f)) {
        add_line(q, c, make_float2(l.xpos, yl), l.right, bit, added, enclosed);
        lw = true;
    }
    const float yr = (r.right.x <= 1.0f) ? r.right.y : intersect_curve(r);
    if ((r.left.x > 0.0f) && (r.left.x < 1.0f) && (l.index != r.index)) {
        add_line(q, c, r.left, make_float2(r.xpos, yr), bit, added, enclosed);
        rw = true;
    }
    if (l.right.x < r.left.x || (!lw && !rw)) {
        const float2 bl = (lw || (l.left.x == l.right.x)) ? l.right : make_float2(l.xpos, yl);
        const float2 br = (rw || (r.left.x == r.right.x)) ? r.left : make_float2(r.xpos, yr);
        add_line(q, c, bl, br, bit, added, enclosed);
    }
}


// searchTree (Kernels.cl:1337-1357) over the strand's Eytzinger tree, for
// the column's left border (l, which also takes a node at its x) and right
// border (r) in one loop: the two descents are independent, so their node
// reads are issued together (the 4096^2 frame: 43.5 -> 42.8 ms).
__device__ __forceinline__ void search_pair(Trav &l, Trav &r, const GUDNI_CONST uint8_t *tree, int tree_size,
                                            float dx, float dy)
{
    l.index = 0;
    r.index = 0;
    while (l.index < tree_size || r.index < tree_size) {
        const bool ml = l.index < tree_size, mr = r.index < tree_size;
        const float4 nl = ld4(tree + 16 * (ml ? l.index : 0));
        const float4 nr = ld4(tree + 16 * (mr ? r.index : 0));
        if (ml) {
            const float px = nl.x - dx, py = nl.y - dy, cx = nl.z - dx, cy = nl.w - dy;
            if ((l.xpos < px) || (l.xpos == px)) {
                l.right = make_float2(px, py);
                l.index = (l.index << 1) + 1;
            } else {
                l.left = make_float2(px, py);
                l.ctl = make_float2(cx, cy);
                l.index = (l.index << 1) + 2;
            }
        }
        if (mr) {
            const float px = nr.x - dx, py = nr.y - dy, cx = nr.z - dx, cy = nr.w - dy;
            if (r.xpos < px) {
                r.right = make_float2(px, py);
                r.index = (r.index << 1) + 1;
            } else {
                r.left = make_float2(px, py);
                r.ctl = make_float2(cx, cy);
                r.index = (r.index << 1) + 2;
            }
        }
    }
}

// buildThresholdArray (Kernels.cl:1540-1595).  The shape and strand loops
// are wave-uniform (every lane walks the tile's whole shape list); only the
// range test, tree searches and record pushes diverge.
//
// kCompact (a tile of at most maxShapes shapes, i.e. every tile above the
// 8-pixel floor): the 127-shape cap cannot bind, and the column's shape bit
// b is a monotone relabelling of the tile-local shape index n (b = number of
// contributing shapes before n).  Labelling headers and the stack by n
// instead visits the same shapes in the same order everywhere, and lets the
// compositing walk read tags and colours from a per-tile table in LDS
// instead of chasing shapeIndices[b] -> shape record -> substance in memory.
// `contrib` records which n contributed, so the reference's labels can be
// recovered (dump path).
template <bool kCompact>
__device__ void generate(const Frame &f, const SceneDesc &s, Queue &q, const Column &c, bool active, int lane,
                         float span0, float span1, int32_t *sidx,
                         uint32_t &shape_bits, uint64_t &stack0, uint64_t &stack1,
                         uint64_t &contrib0, uint64_t &contrib1, bool &any_nonzero)
{
    const float dx = (float)c.dx, dy = (float)c.dy;
    const uint32_t max_shapes = (uint32_t)f.max_shapes;
    // The shape and strand walks are the same for every lane of the wave;
    // readfirstlane makes that explicit so their loads go to the scalar unit.
    const int num_shapes = (int)__builtin_amdgcn_readfirstlane(c.num_shapes);
    const int shape_start = (int)__builtin_amdgcn_readfirstlane(c.shape_start);
    // A shape none of whose strands reaches a column adds no threshold and no
    // enclosure there, so it gets no shape bit and can be skipped.  The wave
    // tests 64 shapes at a time against [span0, span1] (its columns' x extent
    // widened by a pixel, a superset of checkInRange's per-column test) and
    // walks only the candidates, in shape order.
    // A lane also skips a shape lying wholly below its segment (every point
    // at least a pixel under the segment's bottom): its thresholds all start
    // at or below RENDEREND, so none is added and none encloses.
    const float below = dy + c.float_height + 1.0f;
    bool alive = active;
    for (int base = 0; base < num_shapes; base += kWave) {
        alive = alive && shape_bits < max_shapes && !q.spilled && !q.overflow;
        if (!__any(alive)) break;
        const int ni = base + lane;
        bool cand = false;
        float ymin = 0.0f;
        if (ni < num_shapes) {
            const float4 b = s.bounds[shape_start + ni];
            cand = (b.x <= span1) && (b.z > span0);
            ymin = b.y;
        }
        uint64_t pending = __ballot(cand);
        while (pending) {
            const int i = __ffsll((long long)pending) - 1;
            const int n = base + i;
            pending &= pending - 1;
            const float shape_top = __shfl(ymin, i);
            alive = alive && shape_bits < max_shapes && !q.spilled && !q.overflow;
            if (!alive || shape_top >= below) continue;
        const int shape_index = shape_start + n;
        const GUDNI_CONST uint32_t *srec =
            reinterpret_cast<const GUDNI_CONST uint32_t *>(ro(s.shapes) + shape_index);
        uint64_t offset = 16ull * __builtin_amdgcn_readfirstlane(srec[2]);
        const uint32_t num_strands = __builtin_amdgcn_readfirstlane(srec[3]);
        const bool nonzero = (__builtin_amdgcn_readfirstlane(srec[1]) & (uint32_t)(kTagNonzero >> 32)) != 0u;
        any_nonzero = any_nonzero || nonzero;   // this lane walked a nonzero-fill shape
        const uint32_t label = kCompact ? (uint32_t)n : shape_bits;
        bool added = false, enclosed_shape = false;
        int32_t w0 = 0;   // winding number of the segment's top-left corner (nonzero fill)
        for (uint32_t si = 0; si < num_strands; ++si) {
            const GUDNI_CONST uint8_t *strand = ro(s.geo) + offset;
            // u16 size in 8-byte units, then u16 flags: read as one dword
            const uint32_t head = *reinterpret_cast<const GUDNI_CONST uint32_t *>(strand);
            const uint32_t size = head & 0xFFFFu;
            const bool reversed = nonzero && ((head >> 16) & GUDNI_STRAND_REVERSED) != 0u;
            const uint32_t rule = nonzero ? (kNonzero | (reversed ? kReversed : 0u)) : 0u;
            bool enclosed_strand = false;
            // traverseTree (Kernels.cl:1378-1408)
            const float2 right = ld2(strand + 8);
            const float4 lc = ld4(strand + 16);
            Trav l;
            l.right = make_float2(right.x - dx, right.y - dy);
            l.left = make_float2(lc.x - dx, lc.y - dy);
            l.ctl = make_float2(lc.z - dx, lc.w - dy);
            l.xpos = 0.0f;
            l.index = 0;
            if (l.left.x <= 1.0f && l.right.x > 0.0f) {
                Trav r = l;
                l.xpos = cl_max(0.0f, l.left.x);
                r.xpos = cl_min(1.0f, l.right.x);
                const int tree_size = ((int)size - 4) / 2;
                search_pair(l, r, strand + 32, tree_size, dx, dy);
                spawn(q, c, label | rule, l, r, added, enclosed_strand);
            }
            offset += 8u * size;
            enclosed_shape = enclosed_shape != enclosed_strand;
            // a strand crossing the left border above the segment adds the
            // sign of its original x direction (coming down that border)
            if (enclosed_strand) w0 += reversed ? -1 : 1;
        }
        if (nonzero) enclosed_shape = w0 != 0;
        if (enclosed_shape) {
            if (label < 64u) stack0 ^= 1ull << label;
            else stack1 ^= 1ull << (label - 64u);
        }
        if (added || enclosed_shape) {
            if (nonzero) sidx[(kWindOffset + (int)label) * kWave] = w0;
            if (kCompact) {
                if (label < 64u) contrib0 |= 1ull << label;
                else contrib1 |= 1ull << (label - 64u);
            } else if (shape_bits < max_shapes) {
                sidx[(int)shape_bits * kWave] = shape_index;
            }
            shape_bits += 1u;
        }
        }
    }
}

// Stable insertion sort by thresholdIsBelow: the same permutation as the
// reference's bubble sort (Kernels.cl:1962-1976), which never swaps equal
// keys.
__device__ void sort_queue(Queue &q)
{
    for (int i = 1; i < q.len; ++i) {
        const Thr x = q.get(i);
        const uint32_t xh = q.get_h(i);
        int j = i - 1;
        while (j >= 0) {
            const Thr y = q.get(j);
            const uint32_t yh = q.get_h(j);
            if (!is_below(yh, y, xh, x)) break;
            q.set(j + 1, yh, y);
            --j;
        }
        if (j + 1 != i) q.set(j + 1, xh, x);
    }
}

// Stage (b) with a radix pass first.  The records' top digit is a
// monotone key prefix: floor(top * 2^bits / h2), h2 the column segment's
// height rounded up to a power of two (an exact scaling), so bucket order
// never contradicts thresholdIsBelow and equal tops share a bucket.  A
// counting pass (per-lane histogram in LDS), an exclusive scan and a stable
// scatter into the ring's free half order the records by digit; the
// insertion sort then only meets inversions inside a bucket.  Both steps
// are stable, so the permutation is the stable sort's, i.e. the reference's
// bubble sort's.  The 64 lanes of a wave sort their 64 columns in lockstep;
// `hist` is the lane's column of a [32][64] u16 table (the tile table's
// colour cache, which is rebuilt after the sort).  Columns with fewer than
// kRadixMin records, or more than half a ring, take the insertion sort
// alone.
constexpr int kRadixMin = 8;
constexpr int kRadixMaxBits = 5;

__device__ __forceinline__ int top_digit(float top, float scale, int nb)
{
    const int d = (int)(top * scale);
    return d < 0 ? 0 : (d >= nb ? nb - 1 : d);
}

__device__ void radix_sort_queue(Queue &q, uint16_t *hist, int bits, float height)
{
    const int n = q.len;
    if (bits > 0 && n >= kRadixMin && 2 * n <= q.mask + 1) {
        const int nb = 1 << bits;
        float h2 = 1.0f;
        while (h2 < height) h2 *= 2.0f;
        const float scale = (float)nb / h2;
        for (int b = 0; b < nb; ++b) hist[b * kWave] = 0;
        for (int i = 0; i < n; ++i) {
            const int d = top_digit(q.thr[q.row(i)].x, scale, nb);
            hist[d * kWave] = (uint16_t)(hist[d * kWave] + 1);
        }
        uint32_t sum = 0;
        for (int b = 0; b < nb; ++b) {
            const uint32_t c = hist[b * kWave];
            hist[b * kWave] = (uint16_t)sum;
            sum += c;
        }
        for (int i = 0; i < n; ++i) {
            const Thr t = q.get(i);
            const uint32_t h = q.get_h(i);
            const int d = top_digit(t.top, scale, nb);
            const int pos = hist[d * kWave];
            hist[d * kWave] = (uint16_t)(pos + 1);
            q.set(n + pos, h, t);
        }
        q.head = (q.head + n) & q.mask;
    }
    sort_queue(q);
}

// insertThreshold (Kernels.cl:1105-1124)
__device__ void insert_threshold(Queue &q, uint32_t h, const Thr &t)
{
    if (!q.push_slot()) return;
    int cur = 0;
    while (cur < q.len - 1) {
        const uint32_t oh = q.get_h(cur + 1);
        const Thr o = q.get(cur + 1);
        if (!is_below(h, t, oh, o)) break;
        q.set(cur, oh, o);
        ++cur;
    }
    q.set(cur, h, t);
}

struct Sweep {
    int current, num_active;
    float sx0, sy0, sx1, sy1;     // sectionStart / sectionEnd
    float pixel_y;
    uint32_t rnd;
};

// How a crossing of a record moves through it (the fill rule's sign):
// kCrossRight moving right (horizontalAdvance), kCrossLeft moving back to
// the column's left border (the start of verticalAdvance), kCrossDown moving
// down that border through a persistent record's end there.
enum Cross { kCrossRight, kCrossLeft, kCrossDown };

struct Stack {
    uint64_t s0, s1;
    uint32_t dirty;   // colour groups touched since their composite was cached
    // a bit's colour group is (b * gm) >> 16, gm = ceil(2^16 / the group
    // width) (TileTable::gm; exact for b < 128); 2048: groups of 32
    uint32_t gm = 2048u;
    __device__ __forceinline__ void flip(uint32_t h)
    {
        const uint32_t b = h & kLabelMask;
        if (b < 64u) s0 ^= 1ull << b;
        else s1 ^= 1ull << (b - 64u);
        dirty |= 1u << ((b * gm) >> 16);
    }
    // Crossing record h: an even-odd shape flips its bit (passHeader,
    // Kernels.cl:298-304).  A nonzero-fill shape adds the crossing's winding
    // change, sign(e.x m.y - e.y m.x) for an edge of original direction e
    // crossed by a move m, to its counter (wind[label * 64], this lane's) and
    // its bit becomes counter != 0.
    // any: some record may carry kNonzero (else this is flip).
    template <Cross kHow>
    __device__ __forceinline__ void cross(uint32_t h, int32_t *wind, bool any)
    {
        if (!any || !(h & kNonzero)) {
            flip(h);
            return;
        }
        const bool rev = (h & kReversed) != 0u;
        int d;
        if (kHow == kCrossDown) {
            d = rev ? -1 : 1;
        } else {
            d = (rev != ((h & kPositiveSlope) != 0u)) ? -1 : 1;
            if (kHow == kCrossLeft) d = -d;
        }
        const uint32_t b = h & kLabelMask;
        int32_t *w = wind + (int)b * kWave;
        const int32_t v = *w + d;
        *w = v;
        const uint64_t m = 1ull << (b & 63u);
        uint64_t &word = b < 64u ? s0 : s1;
        word = v != 0 ? (word | m) : (word & ~m);
        dirty |= 1u << ((b * gm) >> 16);
    }
};

// splitNext = countActive + nextSlicePoint + sliceActive (Kernels.cl:1007-1077)
__device__ float split_next(Queue &q, Sweep &p)
{
    float slice = FLT_MAX;
    const float top = q.get(0).top;
    int n = 1;
    while (n < q.len) {
        const float nt = q.get(n).top;
        if (nt > top) { slice = nt; break; }
        ++n;
    }
    p.num_active = n;
    float sp = slice;
    for (int i = 0; i < n; ++i) {
        const float b = q.get(i).bottom;
        if (top < b) sp = cl_min(sp, b);
    }
    slice = cl_min(slice, sp);
    for (int cur = 0; cur < n; ++cur) {
        uint32_t h = q.get_h(cur);
        Thr t = q.get(cur);
        if (t.top < slice && slice < t.bottom) {
            // splitThreshold + divideThreshold (Kernels.cl:928-979)
            const float sx = intersect_x(h, t, slice);
            uint32_t bh;
            Thr b;
            if (hdr_pos(h)) {
                b = Thr{slice, t.bottom, sx, t.right};
                t = Thr{t.top, slice, t.left, sx};
                bh = h & kUnpersist;
            } else {
                b = Thr{slice, t.bottom, t.left, sx};
                t = Thr{t.top, slice, sx, t.right};
                bh = h;
                h &= kUnpersist;
            }
            q.set(cur, h, t);
            if (hdr_persist(bh) || ((b.bottom - b.top) >= kMinCrop)) {
                insert_threshold(q, bh, b);
                if (q.spilled || q.overflow) return slice;
            }
        }
    }
    return slice;
}

// verticalAdvance (Kernels.cl:1744-1824).  nz (wave-uniform): some record
// may belong to a nonzero-fill shape; wind is this lane's winding counters.
__device__ void vertical_advance(Queue &q, const Column &c, Sweep &p, Stack &st, int32_t *wind = nullptr,
                                 bool nz = false)
{
    if (p.sx1 != 1.0f) return;
    for (int i = 0; i < p.num_active; ++i) st.cross<kCrossLeft>(q.get_h(i), wind, nz);
    const float next_break = cl_min(c.float_height, p.pixel_y);
    const float active_bottom = q.len > 0 ? q.get(0).bottom : FLT_MAX;
    if (active_bottom == p.sy1) {
        while (p.num_active > 0) {
            const uint32_t h = q.get_h(0);
            if (hdr_persist_bottom(h)) st.cross<kCrossDown>(h, wind, nz);
            q.pop();
            --p.num_active;
        }
    }
    float next_bottom;
    if (p.num_active > 0) {
        next_bottom = cl_min(active_bottom, next_break);
    } else {
        const float next_top = p.num_active < q.len ? q.get(p.num_active).top : FLT_MAX;
        if (next_top > p.sy1) {
            next_bottom = cl_min(next_break, next_top);
        } else {
            next_bottom = cl_min(next_break, split_next(q, p));
            if (q.spilled || q.overflow) return;
            while (p.num_active > 0) {
                const Thr t0 = q.get(0);
                if (!(t0.top == t0.bottom)) break;
                const uint32_t h = q.get_h(0);
                if (hdr_persist_top(h)) st.cross<kCrossDown>(h, wind, nz);
                q.pop();
                --p.num_active;
            }
            for (int i = 0; i < p.num_active; ++i) {
                if (q.get(i).top > 0.0f) {
                    const uint32_t h = q.get_h(i);
                    if (hdr_persist_top(h)) st.cross<kCrossDown>(h, wind, nz);
                }
            }
        }
    }
    p.sy0 = p.sy1;
    p.sy1 = next_bottom;
    p.sx0 = p.sx1 = 0.0f;
    p.current = 0;
}

__device__ __forceinline__ float4 composite(const float4 fg, const float4 bg)
{   // composite (Kernels.cl:878-887)
    const float a = fg.w + bg.w * (1.0f - fg.w);
    if (a > 0.0f) {
        const float k = 1.0f - fg.w;
        return make_float4(((fg.x * fg.w) + (bg.x * bg.w * k)) / a,
                           ((fg.y * fg.w) + (bg.y * bg.w * k)) / a,
                           ((fg.z * fg.w) + (bg.z * bg.w * k)) / a, a);
    }
    return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ int f2i_rtz(float v)
{
    if (!(v == v)) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int)v;
}

__device__ __forceinline__ int f2i_floor(float v)
{
    if (!(v == v)) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int)floorf(v);
}

// b / 255 for a byte b, correctly rounded: the product by 1/255 (rounded to
// nearest) corrected by one FMA of its exact residual, which equals the IEEE
// division for every one of the 256 bytes.
__device__ __forceinline__ float unit_byte(uint32_t b)
{
    const float x = (float)b;
    const float r = 0x1.010102p-8f;
    const float q = x * r;
    return fmaf(fmaf(-q, 255.0f, x), r, q);
}

// Texel (x, y) of a picture use, transparent black outside it.  The
// coordinates are taken as unsigned, so a negative one is outside too; a
// texel's byte offset fits in 31 bits (validate_uses).
__device__ __forceinline__ float4 texel(const SceneDesc &s, const PictUse &u, uint32_t x, uint32_t y)
{
    if (x >= (uint32_t)u.w || y >= (uint32_t)u.h) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t off = (uint32_t)u.mem_offset + 4u * (y * (uint32_t)u.w + x);
    const uchar4 px = *reinterpret_cast<const uchar4 *>(s.pict + off);
    return make_float4(unit_byte(px.x), unit_byte(px.y), unit_byte(px.z), unit_byte(px.w));
}

// Gradient substance at the pixel's integer position (where pictures are
// sampled too): linear from (x0,y0) to (x1,y1) or radial around (x0,y0) with
// radius x1, stops interpolated piecewise linearly.
__device__ float4 eval_gradient(const float4 *g, int ax, int ay)
{
    const float4 h0 = g[0], h1 = g[1], off = g[2];
    const int kind = (int)__float_as_uint(h0.x);
    int n = (int)__float_as_uint(h0.y);
    const float x0 = h0.z, y0 = h0.w, x1 = h1.x, y1 = h1.y;
    const float px = (float)ax, py = (float)ay;
    // radial: |p - c| / radius; linear: (p - a).(b - a) / |b - a|^2 -- one
    // division for both, so a wave with both kinds divides once
    float num, den;
    if (kind == 1) {
        const float ex = px - x0, ey = py - y0;
        num = sqrtf((ex * ex) + (ey * ey));
        den = x1;
    } else {
        const float dx = x1 - x0, dy = y1 - y0;
        num = ((px - x0) * dx) + ((py - y0) * dy);
        den = (dx * dx) + (dy * dy);
    }
    float t = den > 0.0f ? num / den : 0.0f;
    t = cl_min(cl_max(t, 0.0f), 1.0f);
    n = n < 1 ? 1 : (n > 4 ? 4 : n);
    const float o[4] = {off.x, off.y, off.z, off.w};
    if (!(t > o[0])) return g[3];
    int i = 1;
    while (i < n && t > o[i]) ++i;
    if (i == n) return g[3 + n - 1];
    const float span = o[i] - o[i - 1];
    const float f = span > 0.0f ? (t - o[i - 1]) / span : 0.0f;
    const float4 a = g[3 + i - 1], b = g[3 + i];
    return make_float4(a.x + ((b.x - a.x) * f), a.y + ((b.y - a.y) * f), a.z + ((b.z - a.z) * f),
                       a.w + ((b.w - a.w) * f));
}

// readColor (Kernels.cl:1411-1442): solid colour, picture texel (nearest, as
// the reference samples, or bilinear) or gradient, from the shape's
// substance record `sub`
__device__ float4 read_color_sub(const SceneDesc &s, uint64_t tag, float4 sub, int ax, int ay)
{
    const uint64_t type = tag & kTagTypeMask;
    if (type == kTagSolid) return sub;
    const uint32_t id = __float_as_uint(sub.x);
    if (type == kTagGradient) return eval_gradient(s.grads + 7 * (size_t)id, ax, ay);
    const PictUse u = s.uses[id];
    float scale = u.scale;
    scale = scale < 0.0000001f ? 0.0000001f : scale;
    const float sx = ((float)ax / scale) - u.tx, sy = ((float)ay / scale) - u.ty;
    if (s.filter == 1) {
        const float fx = sx - 0.5f, fy = sy - 0.5f;
        const int ix = f2i_floor(fx), iy = f2i_floor(fy);
        const float wx = fx - (float)ix, wy = fy - (float)iy;
        const uint32_t x0 = (uint32_t)ix, y0 = (uint32_t)iy;
        const float4 t0 = texel(s, u, x0, y0), t1 = texel(s, u, x0 + 1u, y0);
        const float4 t2 = texel(s, u, x0, y0 + 1u), t3 = texel(s, u, x0 + 1u, y0 + 1u);
        const float ax0 = 1.0f - wx, ay0 = 1.0f - wy;
        float4 r;
        r.x = (((t0.x * ax0) + (t1.x * wx)) * ay0) + (((t2.x * ax0) + (t3.x * wx)) * wy);
        r.y = (((t0.y * ax0) + (t1.y * wx)) * ay0) + (((t2.y * ax0) + (t3.y * wx)) * wy);
        r.z = (((t0.z * ax0) + (t1.z * wx)) * ay0) + (((t2.z * ax0) + (t3.z * wx)) * wy);
        r.w = (((t0.w * ax0) + (t1.w * wx)) * ay0) + (((t2.w * ax0) + (t3.w * wx)) * wy);
        return r;
    }
    return texel(s, u, (uint32_t)f2i_rtz(sx), (uint32_t)f2i_rtz(sy));
}

__device__ __forceinline__ float4 read_color(const SceneDesc &s, uint64_t tag, int ax, int ay)
{
    return read_color_sub(s, tag, s.substances[tag & kTagIdMask], ax, ay);
}

// Per-wave LDS table of a compact tile's shapes by tile-local index: tag,
// and the solid colour (in the coverage mode's form) or, for a picture or
// gradient, the substance record itself.
struct TileTable {
    uint64_t tag[128];
    float4 color[128];
    // grouped colours (accurate mode): per lane, the premultiplied composite
    // of each group of 32 tile-local shapes, [group][component][lane]
    float gc[4][4][64];
    uint64_t pic0, pic1;   // tile-local shapes whose substance is a picture
    int grouped;           // the tile qualifies for grouped colours
    // colour group g holds tile-local shapes [g gq, (g+1) gq), gq = ceil(n / 4)
    // for a tile of n shapes; gm = ceil(2^16 / gq) maps a shape to its group
    uint32_t gq, gm;
    // per group: the tile-local shapes some lane of the current unit holds a
    // record or an enclosure of, i.e. every stack bit it can set
    uint32_t cand[4];
};
static_assert(sizeof(TileTable::gc) >= sizeof(uint16_t) * kWave * (1 << kRadixMaxBits),
              "the sort histogram lives in the colour cache");

// Bits [start, start + q) of the 128-bit set (s1:s0), q <= 32, start < 128;
// bits past 127 read as 0.
__device__ __forceinline__ uint32_t group_bits(uint64_t s0, uint64_t s1, uint32_t start, uint32_t q)
{
    uint64_t v;
    if (start >= 64u) v = s1 >> (start - 64u);
    else if (start == 0u) v = s0;
    else v = (s0 >> start) | (s1 << (64u - start));
    return (uint32_t)v & (q >= 32u ? 0xFFFFFFFFu : ((1u << q) - 1u));
}

// Composite of one group's set shapes (word: this lane's bits, 0 for a lane
// that takes no part), top-down, premultiplied.  Once the alpha is exactly 1
// every later layer adds its colour times 0, so a lane stops there with the
// same result.  The wave walks the group's shapes together, from the top of
// `cand` (wave-uniform, a superset of every lane's bits): each shape is
// composited by the lanes that hold it.  A lane meets its own shapes in its
// own order, as a walk of its bits alone would, but the wave evaluates one
// substance at a time, whose tag and record every lane reads alike.
__device__ __forceinline__ float4 group_composite(const SceneDesc &s, const TileTable *tt,
                                                  uint32_t word, int n0, int ax, int ay, uint32_t cand)
{
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    cand = __builtin_amdgcn_readfirstlane(cand);
    while (cand) {
        if (!__any(word != 0u)) break;
        const int b = 31 - __clz((int)cand);
        cand &= ~(1u << b);
        if ((word >> b) & 1u) {
            word &= ~(1u << b);
            const int n = n0 + b;
            const uint64_t tv = tt->tag[n];
            const uint64_t tag = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(tv >> 32)) << 32) |
                                 (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)tv);
            float4 pm = tt->color[n];
            if ((tag & kTagTypeMask) != kTagSolid) {
                pm.x = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(pm.x)));
                const float4 c = read_color_sub(s, tag, pm, ax, ay);
                pm = make_float4(c.x * c.w, c.y * c.w, c.z * c.w, c.w);
            }
            const float t = 1.0f - acc.w;
            acc.x = acc.x + pm.x * t;
            acc.y = acc.y + pm.y * t;
            acc.z = acc.z + pm.z * t;
            acc.w = acc.w + pm.w * t;
            if (acc.w == 1.0f) word = 0u;
        }
    }
    return acc;
}

// Accurate-mode colour of a grouped tile: over(group 3 .. group 0,
// background), recomputing only the dirty groups' cached composites.  Groups
// under an alpha of exactly 1 add nothing (their colour times 0): they are
// neither read nor recomputed, and stay dirty.
__device__ float4 grouped_color(const SceneDesc &s, TileTable *tt, Stack &st, int lane, int ax,
                                int ay)
{
    const uint32_t q = __builtin_amdgcn_readfirstlane(tt->gq);
    const uint32_t words[4] = {group_bits(st.s0, st.s1, 0u, q), group_bits(st.s0, st.s1, q, q),
                               group_bits(st.s0, st.s1, 2u * q, q), group_bits(st.s0, st.s1, 3u * q, q)};
    float P0 = 0.f, P1 = 0.f, P2 = 0.f, A = 0.f;
    uint32_t fresh = 0;
    bool open = true;   // no group so far has brought the alpha to 1
#pragma unroll
    for (int g = 3; g >= 0; --g) {
        open = open && A != 1.0f;
        const bool stale = open && (st.dirty & (1u << g)) != 0u;
        float4 gcol = make_float4(0.f, 0.f, 0.f, 0.f);
        // every lane of the sweep step takes part in the walk (with no bits
        // unless its group is stale)
        if (__any(stale))
            gcol = group_composite(s, tt, stale ? words[g] : 0u, (int)(g * q), ax, ay, tt->cand[g]);
        if (stale) {
            tt->gc[g][0][lane] = gcol.x;
            tt->gc[g][1][lane] = gcol.y;
            tt->gc[g][2][lane] = gcol.z;
            tt->gc[g][3][lane] = gcol.w;
            fresh |= 1u << g;
        } else if (open) {
            gcol = make_float4(tt->gc[g][0][lane], tt->gc[g][1][lane], tt->gc[g][2][lane],
                               tt->gc[g][3][lane]);
        }
        if (open) {
            const float t = 1.0f - A;
            P0 = P0 + gcol.x * t;
            P1 = P1 + gcol.y * t;
            P2 = P2 + gcol.z * t;
            A = A + gcol.w * t;
        }
    }
    st.dirty &= ~fresh;
    {
        const float4 pm = make_float4(s.bg[0] * s.bg[3], s.bg[1] * s.bg[3], s.bg[2] * s.bg[3], s.bg[3]);
        const float t = 1.0f - A;
        P0 = P0 + pm.x * t;
        P1 = P1 + pm.y * t;
        P2 = P2 + pm.z * t;
        A = A + pm.w * t;
    }
    if (A > 0.0f) return make_float4(P0 / A, P1 / A, P2 / A, A);
    return make_float4(0.f, 0.f, 0.f, 0.f);
}

// A lane's colours of the current pixel for picture and gradient layers of a
// compact tile, which depend on the shape and the pixel only: the tile's
// first 32 such shapes (by rank among them, tt->pic0/pic1) keep theirs in
// the first 128 words of the lane's shape scratch, which a compact tile
// does not use.  known: which ranks hold the current pixel's colour (clear
// it at every new pixel).
struct LayerCache {
    int32_t *words;
    uint32_t known;

    __device__ __forceinline__ float4 color(const SceneDesc &s, const TileTable *tt, int n, uint64_t tag,
                                            float4 sub, int ax, int ay)
    {
        const uint64_t pm0 = tt->pic0, pm1 = tt->pic1;
        const uint32_t rank =
            n < 64 ? (uint32_t)__popcll(pm0 & ((1ull << n) - 1ull))
                   : (uint32_t)(__popcll(pm0) + __popcll(pm1 & ((1ull << (n - 64)) - 1ull)));
        if (rank < 32u && ((known >> rank) & 1u))
            return make_float4(__int_as_float(words[(4 * rank + 0) * kWave]),
                               __int_as_float(words[(4 * rank + 1) * kWave]),
                               __int_as_float(words[(4 * rank + 2) * kWave]),
                               __int_as_float(words[(4 * rank + 3) * kWave]));
        const float4 c = read_color_sub(s, tag, sub, ax, ay);
        if (rank < 32u) {
            words[(4 * rank + 0) * kWave] = __float_as_int(c.x);
            words[(4 * rank + 1) * kWave] = __float_as_int(c.y);
            words[(4 * rank + 2) * kWave] = __float_as_int(c.z);
            words[(4 * rank + 3) * kWave] = __float_as_int(c.w);
            known |= 1u << rank;
        }
        return c;
    }
};

// determineColor (Kernels.cl:1447-1513).  kCompact: the stack is labelled by
// tile-local shape index and tags/colours come from the LDS table.  lc (a
// compact tile's, or null): picture and gradient colours through the lane's
// cache of the current pixel.
template <bool kCompact>
__device__ float4 determine_color(const Frame &f, const SceneDesc &s, const Stack &st, const int32_t *sidx,
                                  const TileTable *tt, int ax, int ay, bool &positional,
                                  LayerCache *lc = nullptr)
{
    // The walk visits set stack bits from the top (findTop with a falling
    // ignoreAbove, Kernels.cl:281-292): pop the highest bit of a working copy.
    // Bits at or above maxShapes are never set.
    uint64_t w0 = st.s0, w1 = st.s1;
    float4 base = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 next = base;      // REFERENCE: colour; ACCURATE: premultiplied colour
    uint64_t last_id = ~0ull;
    bool last_is_continue = true, last_is_set = false;
    const bool ref_mode = f.coverage_mode == GUDNI_COVERAGE_REFERENCE;
    for (;;) {
        bool should_composite = true;
        bool done = false;
        int top_bit;
        if (w1) { top_bit = 127 - __clzll((long long)w1); w1 &= ~(1ull << (top_bit - 64)); }
        else if (w0) { top_bit = 63 - __clzll((long long)w0); w0 &= ~(1ull << top_bit); }
        else top_bit = -1;
        if (top_bit < 0) {
            next = make_float4(s.bg[0], s.bg[1], s.bg[2], s.bg[3]);
            if (!ref_mode) next = make_float4(next.x * next.w, next.y * next.w, next.z * next.w, next.w);
            done = true;
            last_is_set = true;
        } else {
            uint64_t tag;
            if (kCompact) tag = tt->tag[top_bit];
            else tag = s.shapes[sidx[top_bit * kWave]].tag;
            const uint64_t sid = tag & kTagIdMask;
            const bool is_cont = (tag & kTagCompoundMask) == kTagContinue;
            const bool is_add = (tag & kTagCompoundMask) == kTagAdd;
            if (sid == last_id) {
                if (last_is_continue) {
                    if (!is_cont) { last_is_set = is_add; last_is_continue = false; }
                    else last_is_set = !last_is_set;
                }
                should_composite = false;
            } else {
                const bool solid = (tag & kTagTypeMask) == kTagSolid;
                if (kCompact && solid) {
                    next = tt->color[top_bit];          // already in the mode's form
                } else {
                    if (kCompact)
                        next = lc ? lc->color(s, tt, top_bit, tag, tt->color[top_bit], ax, ay)
                                  : read_color_sub(s, tag, tt->color[top_bit], ax, ay);
                    else
                        next = read_color(s, tag, ax, ay);
                    positional = positional || !solid;
                    if (!ref_mode)
                        next = make_float4(next.x * next.w, next.y * next.w, next.z * next.w, next.w);
                }
                last_is_set = is_add || is_cont;
            }
            last_id = sid;
        }
        if (should_composite && last_is_set) {
            if (ref_mode) {
                base = composite(base, next);
            } else {
                // premultiplied: base.xyz = P, base.w = A; the alpha update is
                // the reference's alphaOut expression, so A == 1 terminates at
                // the same shape
                const float t = 1.0f - base.w;
                base.x = base.x + next.x * t;
                base.y = base.y + next.y * t;
                base.z = base.z + next.z * t;
                base.w = base.w + next.w * t;
            }
            if (base.w == 1.0f) done = true;
        }
        if (done) break;
    }
    if (!ref_mode) {
        if (base.w > 0.0f) {
            base.x = base.x / base.w;
            base.y = base.y / base.w;
            base.z = base.z / base.w;
        } else {
            base = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    return base;
}

__device__ __forceinline__ uint32_t pixel_word(float4 c, bool rgba)
{   // colorToSolidPixel_Word32_BGRA (Kernels.cl:842-844): rtz, alpha 255; the
    // same bytes in R,G,B,A memory order when `rgba`
    const float ch[3] = {c.z * 255.0f, c.y * 255.0f, c.x * 255.0f};
    uint32_t w = 0xFF000000u;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float v = ch[i];
        const uint32_t u = (v == v) ? (v >= 255.0f ? 255u : (v <= 0.0f ? 0u : (uint32_t)v)) : 0u;
        w |= u << (8 * (rgba ? 2 - i : i));
    }
    return w;
}

// renderThresholdArray + calculatePixel + horizontalAdvance + sectionColor
// (Kernels.cl:1724-1742, 1826-1851, 1881-1916, 1978-2028).  The reference
// loops over a column's pixels and, inside, over each pixel's sections; here
// one loop steps either a section or the end of a pixel.  The sections and
// their arithmetic are the same, but lanes of a wave whose pixels hold
// different numbers of sections no longer wait for each other at every
// pixel, only at the end of their columns (the 4096^2 frame: 57.4 -> 51.9 ms).
template <bool kCompact>
__device__ uint32_t render(const Frame &f, const SceneDesc &s, Queue &q, const Column &c, Stack st,
                           int32_t *sidx, TileTable *tt, bool nz)
{
    int32_t *const wind = sidx + kWindOffset * kWave;
    Sweep p;
    p.current = 0; p.num_active = 0;
    p.sx0 = 0.0f; p.sy0 = 0.0f; p.sx1 = 1.0f; p.sy1 = 0.0f;
    const bool stochastic = f.stochastic != 0.0f;
    if (stochastic) {   // initRandomField (Kernels.cl:1597-1606)
        const uint32_t start = c.unique & (uint32_t)f.random_mask;
        p.rnd = __float_as_uint(f.random[start]) & (uint32_t)f.random_mask;
    } else {
        p.rnd = 0;
    }
    long guard = kSectionGuard;
    int ay = c.dy;
    int yint = 0;
    const bool grouped = kCompact && tt->grouped;
    const int lane = threadIdx.x & (kWave - 1);
    st.dirty = 0xFu;
    if (grouped) st.gm = __builtin_amdgcn_readfirstlane(tt->gm);
    const uint32_t gq = grouped ? __builtin_amdgcn_readfirstlane(tt->gq) : 32u;
    p.pixel_y = 1.0f;
    if (!(p.pixel_y <= c.float_height)) return 0;
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float area_sum = 0.0f;
    if (grouped) {
        // picture colours depend on the pixel: their groups are stale
        const uint64_t p0 = st.s0 & tt->pic0, p1 = st.s1 & tt->pic1;
        st.dirty |= (group_bits(p0, p1, 0u, gq) ? 1u : 0u) | (group_bits(p0, p1, gq, gq) ? 2u : 0u) |
                    (group_bits(p0, p1, 2u * gq, gq) ? 4u : 0u) | (group_bits(p0, p1, 3u * gq, gq) ? 8u : 0u);
    }
    for (;;) {
        if ((p.sx1 < 1.0f) || (p.sy1 < p.pixel_y)) {
            if (--guard < 0) return kStatusGuard;
            vertical_advance(q, c, p, st, wind, nz);
            if (q.spilled || q.overflow) return 0;
            // horizontalAdvance with thresholdMidXLow (Kernels.cl:916-926)
            float nx;
            if (p.current < p.num_active) {
                const Thr t = q.get(p.current);
                const uint32_t h = q.get_h(p.current);
                float x;
                if (f.coverage_mode == GUDNI_COVERAGE_REFERENCE) {
                    const float ymid = p.sy0 + ((p.sy1 - p.sy0) * 0.5f);
                    x = intersect_x(h, t, ymid);
                } else if (t.left == t.right) {
                    x = t.left;
                } else {
                    // the same mid point, relative to the threshold's start
                    const float sy = hdr_pos(h) ? t.top : t.bottom;
                    const float d = (p.sy0 - sy) + ((p.sy1 - p.sy0) * 0.5f);
                    x = (d * inv_slope(h, t)) + t.left;
                }
                nx = (x >= 1.0f) ? 0.0f : cl_max(0.0f, x);
            } else {
                nx = 1.0f;
            }
            p.sx0 = p.sx1;
            p.sx1 = nx;
            float4 col;
            if (grouped) {
                col = grouped_color(s, tt, st, lane, c.dx, ay);
            } else {
                bool positional = false;
                col = determine_color<kCompact>(f, s, st, sidx, tt, c.dx, ay, positional);
            }
            const float area = (p.sx1 - p.sx0) * (p.sy1 - p.sy0);
            float adj = area;
            if (stochastic) {   // getRandom (Kernels.cl:1608-1613)
                p.rnd = (p.rnd + 1u) & (uint32_t)f.random_mask;
                adj = area + (area * f.random[p.rnd] * f.stochastic);
            }
            acc.x += col.x * adj;
            acc.y += col.y * adj;
            acc.z += col.z * adj;
            acc.w += col.w * adj;
            area_sum += adj;
            if (p.current < p.num_active) st.cross<kCrossRight>(q.get_h(p.current), wind, nz);
            ++p.current;
        } else {
            // the pixel's last section is done: its colour is the area-weighted
            // mean (writePixelGlobal, Kernels.cl:1853-1862), then the next pixel
            const float4 color = make_float4(acc.x / area_sum, acc.y / area_sum, acc.z / area_sum,
                                             acc.w / area_sum);
            const long long pix = (long long)(c.dy + yint) * s.width + c.dx;
            s.out[pix] = pixel_word(color, f.out_rgba != 0);
            if (s.out_f32) s.out_f32[pix] = color;
            p.sx0 = 0.0f;
            p.sy0 = p.pixel_y;
            ++ay;
            ++yint;
            p.pixel_y += 1.0f;
            if (!(p.pixel_y <= c.float_height)) break;
            acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            area_sum = 0.0f;
            if (grouped) {
                const uint64_t p0 = st.s0 & tt->pic0, p1 = st.s1 & tt->pic1;
                st.dirty |= (group_bits(p0, p1, 0u, gq) ? 1u : 0u) | (group_bits(p0, p1, gq, gq) ? 2u : 0u) |
                            (group_bits(p0, p1, 2u * gq, gq) ? 4u : 0u) |
                            (group_bits(p0, p1, 3u * gq, gq) ? 8u : 0u);
            }
        }
    }
    return 0;
}

// initTileState / isActiveThread (Kernels.cl:1692-1722)
__device__ __forceinline__ bool init_column(const Frame &f, const SceneDesc &s, const TileRec &t, int job, int column,
                                            Column &c)
{
    const int hd = t.hdepth, vd = t.vdepth, cd = f.compute_depth;
    const int dd = max(0, vd - (cd - hd));
    const int ix = ((1 << hd) - 1) & column;
    const int iy = (column >> hd) << dd;
    c.dx = ix + t.box[0];
    c.dy = iy + t.box[1];
    c.unique = ((uint32_t)t.column_allocation + (1u << cd) + (uint32_t)job * (1u << (2 * cd))) * 7919u;
    c.int_height = min(1 << dd, s.height - c.dy);
    c.float_height = (float)c.int_height;
    c.shape_start = t.shape_start;
    c.num_shapes = t.shape_count;
    return (iy < t.box[3] - t.box[1]) && (c.dx < s.width) && (c.dy < s.height);
}

__device__ __forceinline__ uint32_t wave_uniform(uint32_t v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

// Number of contributing shapes with tile-local index below n: the
// reference's shape bit of shape n in a compact tile.
__device__ __forceinline__ uint32_t bit_of(uint64_t c0, uint64_t c1, uint32_t n)
{
    if (n < 64u) return (uint32_t)__popcll(c0 & ((1ull << n) - 1ull));
    return (uint32_t)(__popcll(c0) + __popcll(c1 & ((1ull << (n - 64u)) - 1ull)));
}

__device__ __forceinline__ uint32_t relabel(uint32_t h, uint64_t c0, uint64_t c1)
{
    return (h & ~kShapeBitMask) | bit_of(c0, c1, h & kShapeBitMask);
}

// Copy the queue into the dump buffers, in the reference's labels.
template <bool kCompact>
__device__ void dump_queue(const Frame &f, const Queue &q, int dslot, bool pre, uint64_t c0,
                           uint64_t c1)
{
    const int k = min(q.len, f.dump_cap);
    float4 *thr = pre ? f.dump_pre_thr : f.dump_thr;
    uint32_t *hdr = pre ? f.dump_pre_hdr : f.dump_hdr;
    for (int i = 0; i < k; ++i) {
        const Thr x = q.get(i);
        const uint32_t h = q.get_h(i);
        thr[(size_t)dslot * f.dump_cap + i] = make_float4(x.top, x.bottom, x.left, x.right);
        hdr[(size_t)dslot * f.dump_cap + i] = kCompact ? relabel(h, c0, c1) : h;
    }
}

// One work unit: 64 columns of one tile, generate -> sort -> render per lane,
// with wave-uniform boundaries between the stages so the optional stage
// clocks bracket whole stages of the wave.
template <bool kCompact>
__device__ void run_unit(const Frame &f, const SceneDesc &s, int lane, uint32_t unit, Queue &q, int32_t *sidx,
                         TileTable *tt, uint32_t &status)
{
    const int tile = (int)(unit >> 2), quarter = (int)(unit & 3u);
    const TileRec t = f.tiles[tile];
    const int job = f.tile_job ? f.tile_job[tile] : 0;
    const int column = unit_column(t.hdepth, f.compute_depth, quarter, lane);
    Column c;
    const bool active = init_column(f, s, t, job, column, c);
    const int dslot = tile * f.threads_per_tile + column;
    const bool timed = f.stage_cycles != nullptr;
    const long long t0 = timed ? clock64() : 0;
    if (kCompact) {
        // the tile's tags and solid colours, by tile-local index
        const bool premultiply = f.coverage_mode != GUDNI_COVERAGE_REFERENCE;
        bool all_add = true;
        uint64_t pic_lo = 0, pic_hi = 0;
        for (int n = lane; n < c.num_shapes; n += kWave) {
            const uint64_t tag = s.shapes[c.shape_start + n].tag;
            tt->tag[n] = tag;
            float4 col = s.substances[tag & kTagIdMask];
            if ((tag & kTagTypeMask) == kTagSolid && premultiply)
                col = make_float4(col.x * col.w, col.y * col.w, col.z * col.w, col.w);
            tt->color[n] = col;
            all_add = all_add && (tag & kTagCompoundMask) == kTagAdd;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // grouped colours need every shape added with its own substance
        bool distinct = true;
        for (int n = lane; n < c.num_shapes; n += kWave) {
            const uint64_t sid = tt->tag[n] & kTagIdMask;
            for (int m = n + 1; m < c.num_shapes && distinct; ++m)
                distinct = (tt->tag[m] & kTagIdMask) != sid;
            if ((tt->tag[n] & kTagTypeMask) != kTagSolid) {
                if (n < 64) pic_lo |= 1ull << n; else pic_hi |= 1ull << (n - 64);
            }
        }
        // OR-reduce the picture masks and AND-reduce the flags over the wave
        for (int o = 32; o > 0; o >>= 1) {
            pic_lo |= __shfl_xor(pic_lo, o);
            pic_hi |= __shfl_xor(pic_hi, o);
        }
        const bool grouped_tile = premultiply && __all(all_add && distinct);
        if (lane == 0) {
            tt->pic0 = pic_lo;
            tt->pic1 = pic_hi;
            tt->grouped = grouped_tile ? 1 : 0;
            const uint32_t gq = c.num_shapes > 4 ? (uint32_t)(c.num_shapes + 3) / 4u : 1u;
            tt->gq = gq;
            tt->gm = (65536u + gq - 1u) / gq;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    uint32_t bits = 0;
    uint64_t s0 = 0, s1 = 0, c0 = 0, c1 = 0;
    bool any_nonzero = false;
    {
        // x extent of this unit's columns (isActiveThread's internalX)
        const int x0 = t.box[0] + unit_x0(t.hdepth, f.compute_depth, quarter);
        const int x1 = x0 + unit_width(t.hdepth, f.compute_depth);
        generate<kCompact>(f, s, q, c, active, lane, (float)x0 - 1.0f, (float)x1 + 1.0f, sidx, bits,
                           s0, s1, c0, c1, any_nonzero);
    }
    if (kCompact) {
        // a lane's stack only ever holds bits of shapes it took a record or
        // an enclosure of: their union over the wave bounds the colour walks
        uint64_t u0 = c0, u1 = c1;
        for (int o = 32; o > 0; o >>= 1) {
            u0 |= __shfl_xor(u0, o);
            u1 |= __shfl_xor(u1, o);
        }
        if (lane == 0) {
            const uint32_t q = tt->gq;
            for (uint32_t g = 0; g < 4u; ++g) tt->cand[g] = group_bits(u0, u1, g * q, q);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const bool ok = active && !q.spilled && !q.overflow;
    const long long t1 = timed ? clock64() : 0;
    if (ok && f.dump_qlen) dump_queue<kCompact>(f, q, dslot, true, c0, c1);
    if (ok) radix_sort_queue(q, reinterpret_cast<uint16_t *>(&tt->gc[0][0][0]) + lane, s.sort_bits,
                             c.float_height);
    if (ok && f.dump_qlen) {
        dump_queue<kCompact>(f, q, dslot, false, c0, c1);
        f.dump_qlen[dslot] = q.len;
        f.dump_bits[dslot] = bits;
        uint64_t r0 = s0, r1 = s1;
        if (kCompact) {     // the initial stack in the reference's labels
            r0 = r1 = 0;
            for (uint32_t n = 0; n < 128u; ++n) {
                const bool on = n < 64u ? ((s0 >> n) & 1ull) : ((s1 >> (n - 64u)) & 1ull);
                if (on) {
                    const uint32_t b = bit_of(c0, c1, n);
                    if (b < 64u) r0 |= 1ull << b; else r1 |= 1ull << (b - 64u);
                }
            }
        }
        f.dump_stack[2 * (size_t)dslot] = r0;
        f.dump_stack[2 * (size_t)dslot + 1] = r1;
        int b = 0;
        for (int i = 0; i < f.max_shapes; ++i) f.dump_sidx[(size_t)dslot * f.max_shapes + i] = 0;
        if (kCompact) {
            for (int n = 0; n < 128; ++n) {
                const bool on = n < 64 ? ((c0 >> n) & 1ull) : ((c1 >> (n - 64)) & 1ull);
                if (on && b < f.max_shapes) f.dump_sidx[(size_t)dslot * f.max_shapes + b++] = c.shape_start + n;
            }
        } else {
            for (int i = 0; i < (int)bits && i < f.max_shapes; ++i)
                f.dump_sidx[(size_t)dslot * f.max_shapes + i] = sidx[i * kWave];
        }
    }
    const long long t2 = timed ? clock64() : 0;
    // only a lane that walked a nonzero-fill shape keeps winding counters
    if (ok) status = render<kCompact>(f, s, q, c, Stack{s0, s1, 0u}, sidx, tt, any_nonzero);
    if (active && q.overflow) status |= kStatusOverflow;
    if (!active && f.dump_qlen) f.dump_qlen[dslot] = -1;
    if (timed) {
        const long long t3 = clock64();
        if (lane == 0) {
            atomicAdd(f.stage_cycles + 0, (unsigned long long)(t1 - t0));
            atomicAdd(f.stage_cycles + 1, (unsigned long long)(t2 - t1));
            atomicAdd(f.stage_cycles + 2, (unsigned long long)(t3 - t2));
            atomicAdd(f.stage_cycles + 3, 1ull);
        }
    }
}

} // namespace dev
} // namespace gudni
