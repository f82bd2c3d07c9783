// raster.hip -- gfx950 kernels and C ABI of libgudni_hip.
//
// One lane owns one column segment of a tile (the reference's work-item,
// Kernels.cl:1692-1722) and runs threshold generation, the per-column sort
// and the scanline coverage/compositing sweep back to back, so records never
// leave the wave's scratch between stages.  Work is claimed by whole
// wavefronts: a work unit is (tile, quarter) = 64 consecutive columns of a
// tile, and quarters without an active column are never queued (an 8x8 tile
// is a single unit).  The grid is persistent; every wave loops on an atomic
// work counter until the list is exhausted, so every wave reaches the exit.
//
// Per-lane threshold queues live in a column-interleaved scratch ring
// (element i of lane l at row*64 + l): lanes touching the same row hit one
// contiguous 1 KiB span.  The fast pass gives each lane a ring of kFastCap
// records; a lane whose live queue outgrows it abandons its unit, which is
// re-run by a second persistent pass whose rings hold maxThresholds records.
// The result is identical to one pass at full capacity, with ~1/32 of the
// scratch.
//
// Numerics: compiled with -ffp-contract=off and correctly rounded FP32
// division, and every FP32 expression is evaluated in the order the
// reference writes it, so records, their order and the coverage sums are
// bit-identical to a strict (non-relaxed) evaluation of Kernels.cl.
// min/max are the OpenCL definitions: max(x,y) = x < y ? y : x,
// min(x,y) = y < x ? y : x.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

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
constexpr float kMinCrop = 0.2f;          // MINCROP (Kernels.cl:47)
constexpr float kFlatness = 0.25f;        // TAXICAB_FLATNESS (Kernels.cl:250)

constexpr uint64_t kTagTypeMask = 0xC000000000000000ull;
constexpr uint64_t kTagSolid = 0x8000000000000000ull;
constexpr uint64_t kTagCompoundMask = 0x3000000000000000ull;
constexpr uint64_t kTagContinue = 0x1000000000000000ull;
constexpr uint64_t kTagAdd = 0x2000000000000000ull;
constexpr uint64_t kTagIdMask = 0x0FFFFFFFFFFFFFFFull;

constexpr int kWave = 64;
constexpr int kFastCapLog2 = 7;           // 128-record rings in the fast pass
constexpr long kSectionGuard = 1L << 22;  // sweep-length bound per column: the
                                          // longest legitimate sweep is 256 rows
                                          // x (2*4096+1) sections = 2.1M
constexpr int kMaxBisect = 64;            // bifurcation bound per intersection

// status codes a lane can raise (bit flags in the launch's status word)
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
    float bg[4];                  // background RGBA
    int32_t width, height;
    int32_t pad0, pad1;
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
    // scratch
    float4 *q_thr;                // [slots][cap][64]
    uint32_t *q_hdr;              // [slots][cap][64]
    int32_t *shape_idx;           // [slots][max_shapes][64]
    int32_t cap_log2;
    // control
    uint32_t *counter;            // work counter for this pass
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

// A lane's threshold queue: a ring in the interleaved scratch.  Logical
// element 0 is the front of the reference's ThresholdQueue (qSlice.sStart).
struct Queue {
    float4 *thr;      // lane's column of the slot's ring (stride 64)
    uint32_t *hdr;
    int head, len, mask, maxt;
    bool spilled;     // ring too small: abandon the unit
    bool overflow;    // live queue > maxThresholds (reference would corrupt)

    __device__ __forceinline__ int row(int i) const { return ((head + i) & mask) * kWave; }
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

// searchTree (Kernels.cl:1337-1357) over the strand's Eytzinger tree
__device__ __forceinline__ void search_tree(Trav &tr, const GUDNI_CONST uint8_t *tree, int tree_size,
                                            float dx, float dy, bool is_left)
{
    tr.index = 0;
    while (tr.index < tree_size) {
        const float4 n = ld4(tree + 16 * tr.index);
        const float px = n.x - dx, py = n.y - dy, cx = n.z - dx, cy = n.w - dy;
        if ((tr.xpos < px) || (is_left && tr.xpos == px)) {
            tr.right = make_float2(px, py);
            tr.index = (tr.index << 1) + 1;
        } else {
            tr.left = make_float2(px, py);
            tr.ctl = make_float2(cx, cy);
            tr.index = (tr.index << 1) + 2;
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
                         uint64_t &contrib0, uint64_t &contrib1)
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
        const uint32_t label = kCompact ? (uint32_t)n : shape_bits;
        bool added = false, enclosed_shape = false;
        for (uint32_t si = 0; si < num_strands; ++si) {
            const GUDNI_CONST uint8_t *strand = ro(s.geo) + offset;
            // u16 size in 8-byte units, then u16 0: read as one dword
            const uint32_t size = *reinterpret_cast<const GUDNI_CONST uint32_t *>(strand) & 0xFFFFu;
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
                search_tree(l, strand + 32, tree_size, dx, dy, true);
                search_tree(r, strand + 32, tree_size, dx, dy, false);
                spawn(q, c, label, l, r, added, enclosed_strand);
            }
            offset += 8u * size;
            enclosed_shape = enclosed_shape != enclosed_strand;
        }
        if (enclosed_shape) {
            if (label < 64u) stack0 ^= 1ull << label;
            else stack1 ^= 1ull << (label - 64u);
        }
        if (added || enclosed_shape) {
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

struct Stack {
    uint64_t s0, s1;
    uint32_t dirty;   // groups of 32 bits touched since their composite was cached
    __device__ __forceinline__ void flip(uint32_t h)
    {
        const uint32_t b = h & kShapeBitMask;
        if (b < 64u) s0 ^= 1ull << b;
        else if (b < 128u) s1 ^= 1ull << (b - 64u);
        dirty |= 1u << ((b >> 5) & 3u);
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

// verticalAdvance (Kernels.cl:1744-1824)
__device__ void vertical_advance(Queue &q, const Column &c, Sweep &p, Stack &st)
{
    if (p.sx1 != 1.0f) return;
    for (int i = 0; i < p.num_active; ++i) st.flip(q.get_h(i));
    const float next_break = cl_min(c.float_height, p.pixel_y);
    const float active_bottom = q.len > 0 ? q.get(0).bottom : FLT_MAX;
    if (active_bottom == p.sy1) {
        while (p.num_active > 0) {
            const uint32_t h = q.get_h(0);
            if (hdr_persist_bottom(h)) st.flip(h);
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
                if (hdr_persist_top(h)) st.flip(h);
                q.pop();
                --p.num_active;
            }
            for (int i = 0; i < p.num_active; ++i) {
                if (q.get(i).top > 0.0f) {
                    const uint32_t h = q.get_h(i);
                    if (hdr_persist_top(h)) st.flip(h);
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

// readColor (Kernels.cl:1411-1442): solid colour or nearest texel
__device__ float4 read_color(const SceneDesc &s, uint64_t sid, bool solid, int ax, int ay)
{
    const float4 sub = s.substances[sid];
    if (solid) return sub;
    const uint32_t pict_id = __float_as_uint(sub.x);
    const PictUse u = s.uses[pict_id];
    float scale = u.scale;
    scale = scale < 0.0000001f ? 0.0000001f : scale;
    const int rx = f2i_rtz(((float)ax / scale) - u.tx);
    const int ry = f2i_rtz(((float)ay / scale) - u.ty);
    if (rx >= 0 && ry >= 0 && rx < u.w && ry < u.h) {
        const uchar4 px = *reinterpret_cast<const uchar4 *>(
            s.pict + u.mem_offset + 4ll * ((long long)ry * u.w + rx));
        return make_float4((float)px.x / 255.0f, (float)px.y / 255.0f, (float)px.z / 255.0f,
                           (float)px.w / 255.0f);
    }
    return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Per-wave LDS table of a compact tile's shapes: tag and (solid) colour by
// tile-local index.
struct TileTable {
    uint64_t tag[128];
    float4 color[128];
    // grouped colours (accurate mode): per lane, the premultiplied composite
    // of each group of 32 tile-local shapes, [group][component][lane]
    float gc[4][4][64];
    uint64_t pic0, pic1;   // tile-local shapes whose substance is a picture
    int grouped;           // the tile qualifies for grouped colours
};

// Composite of one group's set shapes, top-down, premultiplied.
__device__ __forceinline__ float4 group_composite(const SceneDesc &s, const TileTable *tt,
                                                  uint32_t word, int g, int ax, int ay)
{
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    while (word) {
        const int b = 31 - __clz((int)word);
        word &= ~(1u << b);
        const int n = 32 * g + b;
        float4 pm;
        const uint64_t tag = tt->tag[n];
        if ((tag & kTagTypeMask) == kTagSolid) {
            pm = tt->color[n];
        } else {
            const float4 c = read_color(s, tag & kTagIdMask, false, ax, ay);
            pm = make_float4(c.x * c.w, c.y * c.w, c.z * c.w, c.w);
        }
        const float t = 1.0f - acc.w;
        acc.x = acc.x + pm.x * t;
        acc.y = acc.y + pm.y * t;
        acc.z = acc.z + pm.z * t;
        acc.w = acc.w + pm.w * t;
    }
    return acc;
}

// Accurate-mode colour of a grouped tile: over(group 3 .. group 0,
// background), recomputing only the dirty groups' cached composites.
__device__ float4 grouped_color(const SceneDesc &s, TileTable *tt, Stack &st, int lane, int ax,
                                int ay)
{
    const uint32_t words[4] = {(uint32_t)st.s0, (uint32_t)(st.s0 >> 32), (uint32_t)st.s1,
                               (uint32_t)(st.s1 >> 32)};
    float P0 = 0.f, P1 = 0.f, P2 = 0.f, A = 0.f;
#pragma unroll
    for (int g = 3; g >= 0; --g) {
        float4 gcol;
        if (st.dirty & (1u << g)) {
            gcol = group_composite(s, tt, words[g], g, ax, ay);
            tt->gc[g][0][lane] = gcol.x;
            tt->gc[g][1][lane] = gcol.y;
            tt->gc[g][2][lane] = gcol.z;
            tt->gc[g][3][lane] = gcol.w;
        } else {
            gcol = make_float4(tt->gc[g][0][lane], tt->gc[g][1][lane], tt->gc[g][2][lane],
                               tt->gc[g][3][lane]);
        }
        const float t = 1.0f - A;
        P0 = P0 + gcol.x * t;
        P1 = P1 + gcol.y * t;
        P2 = P2 + gcol.z * t;
        A = A + gcol.w * t;
    }
    st.dirty = 0;
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

// determineColor (Kernels.cl:1447-1513).  kCompact: the stack is labelled by
// tile-local shape index and tags/colours come from the LDS table.
template <bool kCompact>
__device__ float4 determine_color(const Frame &f, const SceneDesc &s, const Stack &st, const int32_t *sidx,
                                  const TileTable *tt, int ax, int ay, bool &positional)
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
                    next = read_color(s, sid, solid, ax, ay);
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

__device__ __forceinline__ uint32_t bgra_word(float4 c)
{   // colorToSolidPixel_Word32_BGRA (Kernels.cl:842-844): rtz, alpha 255
    const float ch[3] = {c.z * 255.0f, c.y * 255.0f, c.x * 255.0f};
    uint32_t w = 0xFF000000u;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float v = ch[i];
        const uint32_t u = (v == v) ? (v >= 255.0f ? 255u : (v <= 0.0f ? 0u : (uint32_t)v)) : 0u;
        w |= u << (8 * i);
    }
    return w;
}

// renderThresholdArray + calculatePixel + horizontalAdvance + sectionColor
// (Kernels.cl:1724-1742, 1826-1851, 1881-1916, 1978-2028)
template <bool kCompact>
__device__ uint32_t render(const Frame &f, const SceneDesc &s, Queue &q, const Column &c, Stack st,
                           const int32_t *sidx, TileTable *tt)
{
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


    int yint = -1;
    const bool grouped = kCompact && tt->grouped;
    const int lane = threadIdx.x & (kWave - 1);
    st.dirty = 0xFu;
    for (p.pixel_y = 1.0f; p.pixel_y <= c.float_height; p.pixel_y += 1.0f) {
        ++yint;
        float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float area_sum = 0.0f;
        if (grouped) {
            // picture colours depend on the pixel: their groups are stale
            const uint64_t p0 = st.s0 & tt->pic0, p1 = st.s1 & tt->pic1;
            st.dirty |= ((uint32_t)p0 ? 1u : 0u) | ((p0 >> 32) ? 2u : 0u) |
                        ((uint32_t)p1 ? 4u : 0u) | ((p1 >> 32) ? 8u : 0u);
        }
        while ((p.sx1 < 1.0f) || (p.sy1 < p.pixel_y)) {
            if (--guard < 0) return kStatusGuard;
            vertical_advance(q, c, p, st);
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
            if (p.current < p.num_active) st.flip(q.get_h(p.current));
            ++p.current;
        }
        const float4 color = make_float4(acc.x / area_sum, acc.y / area_sum, acc.z / area_sum,
                                         acc.w / area_sum);
        const long long pix = (long long)(c.dy + yint) * s.width + c.dx;
        s.out[pix] = bgra_word(color);
        if (s.out_f32) s.out_f32[pix] = color;
        p.sx0 = 0.0f;
        p.sy0 = p.pixel_y;
        ++ay;
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
    const int column = quarter * kWave + lane;
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
            float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
            if ((tag & kTagTypeMask) == kTagSolid) {
                col = s.substances[tag & kTagIdMask];
                if (premultiply) col = make_float4(col.x * col.w, col.y * col.w, col.z * col.w, col.w);
            }
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
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    uint32_t bits = 0;
    uint64_t s0 = 0, s1 = 0, c0 = 0, c1 = 0;
    {
        // x extent of this unit's columns (isActiveThread's internalX)
        const int hd = t.hdepth;
        const int w = 1 << hd;
        const int x0 = t.box[0] + (hd >= 6 ? ((quarter * kWave) & (w - 1)) : 0);
        const int x1 = x0 + (hd >= 6 ? kWave : w);
        generate<kCompact>(f, s, q, c, active, lane, (float)x0 - 1.0f, (float)x1 + 1.0f, sidx, bits,
                           s0, s1, c0, c1);
    }
    const bool ok = active && !q.spilled && !q.overflow;
    const long long t1 = timed ? clock64() : 0;
    if (ok && f.dump_qlen) dump_queue<kCompact>(f, q, dslot, true, c0, c1);
    if (ok) sort_queue(q);
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
    if (ok) status = render<kCompact>(f, s, q, c, Stack{s0, s1, 0u}, sidx, tt);
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

// One persistent pass.  A wave claims a work unit, its lanes run their
// columns to completion, and the wave loops until the list is empty.
__global__ void __launch_bounds__(256) raster_pass(Frame f, int from_retry)
{
    __shared__ TileTable tables[4];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int slot = blockIdx.x * (blockDim.x / kWave) + wave;
    const int cap = 1 << f.cap_log2;
    float4 *q_thr = f.q_thr + (size_t)slot * cap * kWave + lane;
    uint32_t *q_hdr = f.q_hdr + (size_t)slot * cap * kWave + lane;
    int32_t *sidx = f.shape_idx + (size_t)slot * f.max_shapes * kWave + lane;
    TileTable *tt = &tables[wave];
    const int n_work = from_retry ? (int)min((uint32_t)f.retry_capacity, f.retry[0]) : f.n_work;
    for (;;) {
        uint32_t w = 0;
        if (lane == 0) w = atomicAdd(f.counter, 1u);
        w = wave_uniform(w);
        if ((int)w >= n_work) break;
        const uint32_t unit = from_retry ? f.retry[1 + w] : f.work[w];
        const int count = f.tiles[unit >> 2].shape_count;   // wave-uniform
        const int scene = f.tile_scene
                              ? (int)__builtin_amdgcn_readfirstlane((uint32_t)f.tile_scene[unit >> 2])
                              : 0;
        SceneDesc s;                                  // wave-uniform: scalar loads
        {
            const GUDNI_CONST uint64_t *src =
                reinterpret_cast<const GUDNI_CONST uint64_t *>(ro(f.scenes + scene));
            uint64_t *dst = reinterpret_cast<uint64_t *>(&s);
#pragma unroll
            for (int i = 0; i < (int)(sizeof(SceneDesc) / 8); ++i) dst[i] = src[i];
        }
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
// Host side of the device API
// ===========================================================================
using namespace gudni;

namespace {

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

} // namespace

struct gudni_ctx {
    gudni_spec spec;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
    int cus = 0;
    int fast_blocks = 0, slow_blocks = 0;
    // common buffers
    DevBuf geo, subs, pict, uses, random;
    int32_t n_substances = 0, n_uses = 0;
    // host copies of the common buffers, used to bounds-check job records
    std::vector<uint8_t> geo_host;
    std::vector<float4> subs_host;
    // per launch buffers
    DevBuf tiles, tile_job, shapes, xrange, work, out, out_f32;
    DevBuf fast_thr, fast_hdr, fast_sidx, slow_thr, slow_hdr, slow_sidx;
    DevBuf control;   // [0] fast counter [1] slow counter [2] status [3..] retry
    DevBuf scene_table;           // single-frame scene descriptor
    dev::SceneDesc single_scene;  // host copy of what scene_table holds
    DevBuf cycles;    // stage cycle counters of the fast pass (timing only)
    DevBuf dump;      // dump staging
    int32_t retry_capacity = 0;
    // uploaded frame (Level 2)
    int32_t frame_w = 0, frame_h = 0, frame_n_work = 0, frame_n_tiles = 0, frame_index = 0;
    float background[4] = {1, 1, 1, 1};
    bool frame_ready = false;
    // timing
    bool timing = false, timed_render = false;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    float last_ms[2] = {0, 0};
    unsigned long long last_cycles[4] = {0, 0, 0, 0};
};

#define CHECK_HIP(ctx, expr)                                                        \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            (ctx)->error = std::string(#expr) + ": " + hipGetErrorString(e_);      \
            return GUDNI_ERR_HIP;                                                   \
        }                                                                           \
    } while (0)

static int set_error(gudni_ctx *ctx, int code, const std::string &msg)
{
    ctx->error = msg;
    return code;
}

// Which quarters of a tile have an active column (isActiveThread,
// Kernels.cl:1720-1722), evaluated for every column on the host.
static void build_work(const gudni_spec &spec, const TileRec *tiles, int n_tiles, int width,
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
                const int col = qd * 64 + l;
                const int ix = ((1 << hd) - 1) & col, iy = (col >> hd) << dd;
                any = iy < (r.box[3] - r.box[1]) && ix + r.box[0] < width && iy + r.box[1] < height;
            }
            if (any) work.push_back(((uint32_t)t << 2) | (uint32_t)qd);
        }
    }
}

static int ensure_scratch(gudni_ctx *ctx)
{
    if (ctx->fast_blocks) return GUDNI_OK;
    const int waves_per_block = 4;
    int per_cu = 0;
    CHECK_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dev::raster_pass, 256, 0));
    if (per_cu < 1) per_cu = 1;
    ctx->fast_blocks = ctx->cus * per_cu;
    ctx->fast_blocks = std::min(ctx->fast_blocks, ctx->cus * 4);
    ctx->slow_blocks = std::max(1, ctx->cus / 8);
    const size_t fast_slots = (size_t)ctx->fast_blocks * waves_per_block;
    const size_t slow_slots = (size_t)ctx->slow_blocks * waves_per_block;
    const size_t fcap = (size_t)1 << dev::kFastCapLog2;
    int slog = 0;
    while ((1 << slog) < ctx->spec.max_thresholds) ++slog;
    const size_t scap = (size_t)1 << slog;
    const size_t ms = (size_t)ctx->spec.max_shapes;
    CHECK_HIP(ctx, ctx->fast_thr.ensure(fast_slots * fcap * 64 * 16));
    CHECK_HIP(ctx, ctx->fast_hdr.ensure(fast_slots * fcap * 64 * 4));
    CHECK_HIP(ctx, ctx->fast_sidx.ensure(fast_slots * ms * 64 * 4));
    CHECK_HIP(ctx, ctx->slow_thr.ensure(slow_slots * scap * 64 * 16));
    CHECK_HIP(ctx, ctx->slow_hdr.ensure(slow_slots * scap * 64 * 4));
    CHECK_HIP(ctx, ctx->slow_sidx.ensure(slow_slots * ms * 64 * 4));
    return GUDNI_OK;
}

static int ensure_control(gudni_ctx *ctx, int n_work)
{
    const int want = std::max(1024, n_work);
    if (ctx->retry_capacity < want) {
        CHECK_HIP(ctx, ctx->control.ensure(sizeof(uint32_t) * (size_t)(4 + want)));
        ctx->retry_capacity = want;
    }
    return GUDNI_OK;
}

// Fill a Frame with the context's buffers; caller sets the per-launch parts.
static dev::Frame base_frame(gudni_ctx *ctx)
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
    for (int i = 0; i < 4; ++i) sd.bg[i] = bg[i];
    sd.width = w;
    sd.height = h;
    return sd;
}

// Both passes, stream-ordered.  The retry pass runs unconditionally; with no
// abandoned units its waves exit on their first claim.
static int launch_passes(gudni_ctx *ctx, dev::Frame f)
{
    uint32_t *ctl = (uint32_t *)ctx->control.p;
    hipLaunchKernelGGL(dev::reset_counters, dim3(1), dim3(64), 0, ctx->stream, ctl, 4);
    CHECK_HIP(ctx, hipGetLastError());
    f.status = ctl + 2;
    f.retry = ctl + 3;
    f.retry_capacity = ctx->retry_capacity;
    // fast pass: small rings, as many waves as fit
    f.counter = ctl + 0;
    f.q_thr = (float4 *)ctx->fast_thr.p;
    f.q_hdr = (uint32_t *)ctx->fast_hdr.p;
    f.shape_idx = (int32_t *)ctx->fast_sidx.p;
    f.cap_log2 = dev::kFastCapLog2;
    const int fast_grid = std::max(1, std::min(ctx->fast_blocks, (f.n_work + 3) / 4));
    hipLaunchKernelGGL(dev::raster_pass, dim3(fast_grid), dim3(256), 0, ctx->stream, f, 0);
    CHECK_HIP(ctx, hipGetLastError());
    if (ctx->timed_render) CHECK_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    // retry pass: rings of maxThresholds records
    int slog = 0;
    while ((1 << slog) < ctx->spec.max_thresholds) ++slog;
    f.counter = ctl + 1;
    f.q_thr = (float4 *)ctx->slow_thr.p;
    f.q_hdr = (uint32_t *)ctx->slow_hdr.p;
    f.shape_idx = (int32_t *)ctx->slow_sidx.p;
    f.cap_log2 = slog;
    hipLaunchKernelGGL(dev::raster_pass, dim3(ctx->slow_blocks), dim3(256), 0, ctx->stream, f, 1);
    CHECK_HIP(ctx, hipGetLastError());
    return GUDNI_OK;
}

static int read_status(gudni_ctx *ctx)
{
    uint32_t st = 0;
    CHECK_HIP(ctx, hipMemcpyAsync(&st, (uint32_t *)ctx->control.p + 2, 4, hipMemcpyDeviceToHost,
                                  ctx->stream));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (st & dev::kStatusOverflow)
        return set_error(ctx, GUDNI_ERR_OVERFLOW, "a column's threshold queue exceeded maxThresholds");
    if (st & dev::kStatusGuard)
        return set_error(ctx, GUDNI_ERR_GUARD, "a column exceeded the section-count guard");
    return GUDNI_OK;
}

template <typename T>
static hipError_t upload(gudni_ctx *ctx, DevBuf &b, const T *src, size_t bytes)
{
    hipError_t e = b.ensure(bytes);
    if (e != hipSuccess) return e;
    if (bytes && src) return hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    return hipSuccess;
}

// Host views of a scene's common buffers, for validation.
struct CommonView {
    const uint8_t *geo;
    size_t geo_bytes;
    const float4 *subs;
    int32_t n_subs, n_uses;
};

static CommonView ctx_view(gudni_ctx *ctx)
{
    return CommonView{ctx->geo_host.data(), ctx->geo_host.size(), ctx->subs_host.data(),
                      ctx->n_substances, ctx->n_uses};
}

static int validate_uses(gudni_ctx *ctx, const void *pict_uses, int32_t n_pict_uses,
                         size_t pict_bytes)
{
    for (int32_t i = 0; i < n_pict_uses; ++i) {
        PictUse u;
        std::memcpy(&u, (const uint8_t *)pict_uses + 24 * (size_t)i, 24);
        const int64_t need = (int64_t)u.mem_offset + 4 * (int64_t)u.w * (int64_t)u.h;
        if (u.w < 0 || u.h < 0 || u.mem_offset < 0 || (u.mem_offset & 3) || need > (int64_t)pict_bytes)
            return set_error(ctx, GUDNI_ERR_ARG, "picture usage " + std::to_string(i) +
                                                     " outside the picture data");
    }
    return GUDNI_OK;
}

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
    if (s.threads_per_tile != 256 || s.max_shapes < 1 || s.max_shapes > 127 ||
        s.max_thresholds < 1 || s.random_field_size < 1 ||
        (s.coverage_mode != GUDNI_COVERAGE_ACCURATE && s.coverage_mode != GUDNI_COVERAGE_REFERENCE) ||
        (s.random_field_size & (s.random_field_size - 1))) {
        g_host_error = "gudni_create: unsupported spec (threads_per_tile must be 256, "
                       "max_shapes 1..127, random_field_size a power of two)";
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
    *out = ctx;
    return GUDNI_OK;
}

void gudni_destroy(gudni_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    DevBuf *bufs[] = {&ctx->geo, &ctx->subs, &ctx->pict, &ctx->uses, &ctx->random,
                      &ctx->tiles, &ctx->tile_job, &ctx->shapes, &ctx->xrange, &ctx->work, &ctx->out,
                      &ctx->out_f32, &ctx->fast_thr, &ctx->fast_hdr, &ctx->fast_sidx,
                      &ctx->slow_thr, &ctx->slow_hdr, &ctx->slow_sidx, &ctx->control, &ctx->dump,
                      &ctx->cycles, &ctx->scene_table};
    for (DevBuf *b : bufs) b->release();
    for (hipEvent_t &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
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

// Host-side validation of a job's records against the common buffers, so no
// kernel ever reads or writes outside them: tile slices inside the job's
// shapes, pixel origins non-negative, every referenced strand walk inside the
// geometry heap with even 8-byte sizes >= 4 (so float4 reads stay aligned),
// substance ids inside the substance table and picture ids inside the usage
// table.
static int validate_job(gudni_ctx *ctx, const CommonView &cv, const ShapeRec *shapes,
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
    const size_t geo_bytes = cv.geo_bytes;
    for (int32_t s = 0; s < n_shapes; ++s) {
        const ShapeRec &r = shapes[s];
        const uint64_t sid = r.tag & dev::kTagIdMask;
        if (sid >= (uint64_t)cv.n_subs)
            return set_error(ctx, GUDNI_ERR_ARG, "shape substance id outside the substance table");
        if ((r.tag & dev::kTagTypeMask) != dev::kTagSolid) {
            const uint32_t pid = __builtin_bit_cast(uint32_t, cv.subs[sid].x);
            if (pid >= (uint32_t)cv.n_uses)
                return set_error(ctx, GUDNI_ERR_ARG, "picture substance refers to a missing usage");
        }
        size_t pos = 16u * (size_t)r.geo_start;
        // Extent of the shape's strands: x from the end points, y over every
        // point (curves stay inside their control hull).  NaN coordinates can
        // never pass the kernel's range tests, so they are left out; a NaN y
        // makes the y extent unknown, which disables the y test.
        float xlo = INFINITY, xhi = -INFINITY, ylo = INFINITY, yhi = -INFINITY;
        bool y_nan = false;
        for (uint32_t k = 0; k < r.num_strands; ++k) {
            if (pos + 8 > geo_bytes)
                return set_error(ctx, GUDNI_ERR_ARG, "strand header outside the geometry heap");
            uint16_t size;
            std::memcpy(&size, cv.geo + pos, 2);
            if (size < 4 || (size & 1) || pos + 8u * size > geo_bytes)
                return set_error(ctx, GUDNI_ERR_ARG, "malformed strand in the geometry heap");
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
        xrange[4 * (size_t)s + 0] = xlo;
        xrange[4 * (size_t)s + 1] = ylo;
        xrange[4 * (size_t)s + 2] = xhi;
        xrange[4 * (size_t)s + 3] = yhi;
    }
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
    if (host_out)
        CHECK_HIP(ctx, hipMemcpyAsync(ctx->out.p, host_out, 4 * npix, hipMemcpyHostToDevice, ctx->stream));
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
    if (host_out)
        CHECK_HIP(ctx, hipMemcpyAsync(host_out, ctx->out.p, 4 * npix, hipMemcpyDeviceToHost, ctx->stream));
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
    int rc0 = set_single_scene(ctx, single_desc(ctx, ctx->background, ctx->frame_w, ctx->frame_h,
                                                out_dev, out_f32_dev), f);
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
    for (int i = 0; i < n && i < 4; ++i) cycles[i] = ctx->last_cycles[i];
    return GUDNI_OK;
}


// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------
struct gudni_batch {
    int32_t n_scenes = 0;
    int device = 0;
    DevBuf geo, subs, pict, uses, shapes, bounds, tiles, tile_job, tile_scene, work, scenes, done,
        frames, random;
    std::vector<size_t> frame_off;
    std::vector<int32_t> width, height;
    std::vector<int32_t> work_begin, work_count;
    int32_t n_work = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *captured_control = nullptr;
    hipStream_t capture_stream = nullptr;   // capture needs a stream other than the null stream
};

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

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
    std::vector<size_t> geo_off, subs_off, pict_off, uses_off, refs_off, frame_off;
    size_t frame_words = 0;
    for (int32_t s = 0; s < n_scenes; ++s) {
        const gudni_batch_scene &sc = scenes[s];
        if (!sc.jobs || sc.width <= 0 || sc.height <= 0 || sc.n_substances < 0 ||
            sc.n_pict_uses < 0 || (sc.geo_bytes && !sc.geo) || (sc.n_substances && !sc.substances) ||
            (sc.pict_bytes && !sc.pict_data) || (sc.n_pict_uses && !sc.pict_uses))
            return set_error(ctx, GUDNI_ERR_ARG, "gudni_batch_upload: bad scene " + std::to_string(s));
        int rc = validate_uses(ctx, sc.pict_uses, sc.n_pict_uses, sc.pict_bytes);
        if (rc) return rc;
        gudni_jobs_view v;
        gudni_jobs_view_get(sc.jobs, &v);
        std::vector<float4> sub_host((size_t)sc.n_substances);
        if (sc.n_substances) std::memcpy(sub_host.data(), sc.substances, 16u * (size_t)sc.n_substances);
        const CommonView cv{(const uint8_t *)sc.geo, sc.geo_bytes, sub_host.data(), sc.n_substances,
                            sc.n_pict_uses};
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
    if (e == hipSuccess) e = up(b->random, random_field, 4 * (size_t)n_random);
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

static dev::Frame batch_frame(gudni_ctx *ctx, gudni_batch *b)
{
    dev::Frame f = base_frame(ctx);
    f.scenes = (const dev::SceneDesc *)b->scenes.p;
    f.tile_scene = (const int32_t *)b->tile_scene.p;
    f.scene_done = (unsigned long long *)b->done.p + 1;
    f.tiles = (const TileRec *)b->tiles.p;
    f.tile_job = (const int32_t *)b->tile_job.p;
    f.work = (const uint32_t *)b->work.p;
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
    if (b->exec && b->captured_control != ctx->control.p) {
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
        b->captured_control = ctx->control.p;
    }
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

void gudni_batch_free(gudni_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    DevBuf *bufs[] = {&b->geo, &b->subs, &b->pict, &b->uses, &b->shapes, &b->bounds, &b->tiles,
                      &b->tile_job, &b->tile_scene, &b->work, &b->scenes, &b->done, &b->frames,
                      &b->random};
    for (DevBuf *d : bufs) d->release();
    delete b;
}

} // extern "C"
