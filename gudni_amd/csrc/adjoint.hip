// adjoint.hip -- gradients of an image-space L2 loss through the raster path.
//
// The differentiable forward is the raster path in accurate arithmetic:
// per column, thresholds are generated from the strands, sorted and swept
// into sections; a pixel's colour is the area-weighted mean of its sections'
// colours, each the premultiplied front-to-back composite of the shapes set
// in the section's stack.  The loss is L = sum over pixels and RGBA of
// (colour - target)^2.
//
// Derivatives.  With every discrete decision held (which strands a column
// meets, curve subdivision, record drops, the section structure), a pixel's
// accumulated colour sum_j area_j C_j depends on
//  * the section colours C_j, through the substances' RGBA;
//  * the section boundaries: section j ends at the mid-height x of threshold
//    j, x(ym) on the line through the threshold's two generated end points A
//    and B.  Moving x by dx moves area dy*dx from section j+1 to section j, so
//    dL/dx = G . (C_j - C_{j+1}) dy with G = dL/d(accumulated colour), and
//    dx(ym) = (1-u) dA.x + u dB.x - k((1-u) dA.y + u dB.y), u = (ym-A.y)/(B.y-A.y),
//    k = (B.x-A.x)/(B.y-A.y).
//  * the span boundaries: a span of a pixel row ends where a record starts
//    or ends; when that is a line end point, moving it by dy moves area
//    between the span above and the span below: dL/dy = G . (sum_j C_j w_j)
//    evaluated at the boundary for the upper span minus the same for the
//    lower span (zero where the records close up; not zero where a short
//    record was dropped or where records cross inside a span);
// The end points are curve points P(t) (Bernstein weights of the strand
// curve's three points) or a column-border intercept of a chord between two
// curve points (found by the same bisection the forward runs); strand curves
// map to the outline's points through the knob split, whose parameter is
// differentiated through the analytic vertical tangent
// t* = (p0.x - c.x) / (p0.x - 2c.x + p1.x).
//
// One persistent launch does, per work unit (64 columns of a tile):
//   A  generate, sort, forward sweep: the image, the loss and G per pixel;
//   B  generate and sort again (recomputation: no coverage is stored), then a
//      backward sweep that accumulates, per generated line, sum w and
//      sum w u of its boundary weights, and per shape the colour gradient
//      (LDS atomics for a tile's shapes);
//   C  generate once more, mapping each line's weights to its end points and
//      from there to the outline points (global atomics);
//   D  the tile's colour gradients go to the substances (global atomics).
// A record's origin (its generated line) travels in header bits 7..28.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"
#include "device.h"
#include "serialize.h"

namespace gudni {
namespace adj {
using namespace dev;

// In-order position of each slot of a strand's Eytzinger tree, by curve count.
struct InOrder {
    int8_t v[kMaxStrandCurves + 1][kMaxStrandCurves];
    constexpr InOrder() : v{}
    {
        for (int m = 1; m <= kMaxStrandCurves; ++m) {
            int next = 0;
            fill(m - 1, 0, next, v[m]);
        }
    }
    static constexpr void fill(int n, int slot, int &next, int8_t *out)
    {
        if (slot >= n) return;
        fill(n, 2 * slot + 1, next, out);
        out[slot] = (int8_t)next++;
        fill(n, 2 * slot + 2, next, out);
    }
};
__constant__ InOrder kInOrder = InOrder();

struct alignas(16) AdjDesc {
    SceneDesc s;                  // forward buffers; s.out_f32 receives the image
    const CurveSrc *curves;       // per strand curve (strand header dword 1 = first index)
    const float2 *points;         // outline points (the parameters)
    const float4 *target;
    float4 *gpix;                 // per pixel: dL/d(accumulated colour)
    float *d_points;              // 2 floats per point
    float *d_rgba;                // 4 floats per substance
    double *loss;
    int32_t n_points, pad;
};
static_assert(sizeof(AdjDesc) % 8 == 0, "AdjDesc is copied in 8-byte words");

struct AdjFrame {
    Frame f;
    const AdjDesc *scenes;
    float4 *q_orig;               // [slots][2 cap][64]: per generated line k, row 2k holds
                                  // (A.y, B.y, sum w, sum w u), row 2k+1 (dL/dA.y, dL/dB.y, 0, 0)
    // rings the forward sweep runs on ([slots][64][cap], as the threshold
    // rings), so that the sorted records stay for the backward sweep; null:
    // the backward sweep regenerates them
    float4 *copy_thr;
    uint32_t *copy_hdr;
    // [slots][4 cap][64]: per generated line k, rows 4k..4k+3 record what
    // maps its weights to outline points (kk, and per end the curve, kind,
    // ta, tb and the chord's slope and position); null: phase C regenerates
    float4 *q_ends;
};

enum Mode { kRecord = 0, kApply = 1 };

// An end point of a generated line in terms of its strand curve.
struct EndSrc {
    int curve;        // global strand curve index
    int kind;         // 0: P(ta); 1: y of the chord P(ta)P(tb) at x = X; 2: y of P(ta) only
    float ta, tb, X;
};

struct Cut { float y, ta, tb; bool vertical, min_at_b; };

// intersect_curve with the parameter interval of the final chord.
__device__ Cut intersect_track(Trav t)
{
    float ta = 0.0f, tb = 1.0f;
    if (!(t.left.x == t.ctl.x && t.left.y == t.ctl.y)) {
        for (int it = 0; it < kMaxBisect; ++it) {
            const float2 lm = mid_point(t.left, t.ctl);
            const float2 rm = mid_point(t.ctl, t.right);
            const float2 on = mid_point(lm, rm);
            const float flat = fabsf(t.ctl.x - on.x) + fabsf(t.ctl.y - on.y);
            if (!(flat > kFlatness)) break;
            const float tm = 0.5f * (ta + tb);
            if (t.xpos < on.x) { t.ctl = lm; t.right = on; tb = tm; }
            else { t.left = on; t.ctl = rm; ta = tm; }
        }
    }
    Cut c;
    c.ta = ta;
    c.tb = tb;
    c.vertical = t.left.x == t.right.x;
    c.min_at_b = t.right.y < t.left.y;
    c.y = c.vertical ? cl_min(t.left.y, t.right.y)
                     : (((t.right.y - t.left.y) / (t.right.x - t.left.x)) * (t.xpos - t.left.x)) + t.left.y;
    return c;
}


// searchTree (Kernels.cl:1337-1357), also giving the in-order index of the
// curve found, for the column's left border (l, which also takes a node at
// its x) and right border (r) in one loop, their node reads issued together;
// the curves' indices go to lcur and rcur.
__device__ __forceinline__ void search_curve_pair(Trav &l, Trav &r, const GUDNI_CONST uint8_t *tree, int tree_size,
                                                  float dx, float dy, int &lcur, int &rcur)
{
    l.index = 0;
    r.index = 0;
    int llast = -1, rlast = -1;
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
                llast = l.index;
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
                rlast = r.index;
                r.index = (r.index << 1) + 2;
            }
        }
    }
    lcur = llast < 0 ? 0 : 1 + (int)kInOrder.v[tree_size + 1][llast];
    rcur = rlast < 0 ? 0 : 1 + (int)kInOrder.v[tree_size + 1][rlast];
}

__device__ __forceinline__ float2 bez(const Trav &t, float s)
{
    const float a = (1.0f - s) * (1.0f - s), b = 2.0f * s * (1.0f - s), c = s * s;
    return make_float2(a * t.left.x + b * t.ctl.x + c * t.right.x, a * t.left.y + b * t.ctl.y + c * t.right.y);
}

__device__ __forceinline__ void atomic_add2(float *p, float2 g)
{
    if (g.x != 0.0f) atomicAdd(p, g.x);
    if (g.y != 0.0f) atomicAdd(p + 1, g.y);
}

// Gradient on a strand curve's (left, control, right) points -> outline points.
__device__ void chain_curve(const AdjDesc &d, int curve, float2 gl, float2 gc, float2 gr)
{
    const CurveSrc cs = d.curves[curve];
    float2 g0 = gl, g1 = gr;
    if (cs.flags & kCurveReversed) { g0 = gr; g1 = gl; }
    const float2 P0 = d.points[cs.p0], C = d.points[cs.c], P1 = d.points[cs.p1];
    float2 d0, dc, d1;
    float gt = 0.0f;
    const float t = cs.t, s = 1.0f - t;
    const float2 dbdt = make_float2(2.0f * (s * (C.x - P0.x) + t * (P1.x - C.x)),
                                    2.0f * (s * (C.y - P0.y) + t * (P1.y - C.y)));
    if (cs.flags & kCurveLowerHalf) {
        // (P0, (1-t)P0 + tC, P(t))
        d0 = make_float2(g0.x + s * gc.x + s * s * g1.x, g0.y + s * gc.y + s * s * g1.y);
        dc = make_float2(t * gc.x + 2.0f * t * s * g1.x, t * gc.y + 2.0f * t * s * g1.y);
        d1 = make_float2(t * t * g1.x, t * t * g1.y);
        gt = gc.x * (C.x - P0.x) + gc.y * (C.y - P0.y) + g1.x * dbdt.x + g1.y * dbdt.y;
    } else if (cs.flags & kCurveUpperHalf) {
        // (P(t), (1-t)C + tP1, P1)
        d0 = make_float2(s * s * g0.x, s * s * g0.y);
        dc = make_float2(2.0f * t * s * g0.x + s * gc.x, 2.0f * t * s * g0.y + s * gc.y);
        d1 = make_float2(t * t * g0.x + t * gc.x + g1.x, t * t * g0.y + t * gc.y + g1.y);
        gt = g0.x * dbdt.x + g0.y * dbdt.y + gc.x * (P1.x - C.x) + gc.y * (P1.y - C.y);
    } else {
        d0 = g0; dc = gc; d1 = g1;
    }
    if (gt != 0.0f) {
        // t* = (P0.x - C.x) / (P0.x - 2 C.x + P1.x)
        const float den = P0.x - 2.0f * C.x + P1.x;
        if (den != 0.0f) {
            const float k = gt / (den * den);
            d0.x += k * (P1.x - C.x);
            dc.x += k * (P0.x - P1.x);
            d1.x += k * (C.x - P0.x);
        }
    }
    atomic_add2(d.d_points + 2 * cs.p0, d0);
    atomic_add2(d.d_points + 2 * cs.c, dc);
    atomic_add2(d.d_points + 2 * cs.p1, d1);
}

__device__ __forceinline__ void chain_point(const AdjDesc &d, int curve, float t, float2 g)
{
    const float a = (1.0f - t) * (1.0f - t), b = 2.0f * t * (1.0f - t), c = t * t;
    chain_curve(d, curve, make_float2(a * g.x, a * g.y), make_float2(b * g.x, b * g.y),
                make_float2(c * g.x, c * g.y));
}

// Gradient (gx, gy) on a generated line's end point -> outline points.
__device__ void apply_end(const AdjDesc &d, const EndSrc &e, const Trav &tv, float gx, float gy)
{
    if (e.kind == 0) {
        chain_point(d, e.curve, e.ta, make_float2(gx, gy));
    } else if (e.kind == 2) {
        chain_point(d, e.curve, e.ta, make_float2(0.0f, gy));
    } else {
        // y = ya + (X - xa) (yb - ya) / (xb - xa); x is the column border
        const float2 pa = bez(tv, e.ta), pb = bez(tv, e.tb);
        const float dxab = pb.x - pa.x;
        const float v = (e.X - pa.x) / dxab, sl = (pb.y - pa.y) / dxab;
        chain_point(d, e.curve, e.ta, make_float2(-gy * sl * (1.0f - v), gy * (1.0f - v)));
        chain_point(d, e.curve, e.tb, make_float2(-gy * sl * v, gy * v));
    }
}

// addLineSegment + addThreshold with the line's origin: kRecord pushes the
// record with origin bits and notes (A.y, 1/(B.y-A.y)); kApply pushes
// nothing and maps the line's accumulated weights to its end points.
// The chord of a kind-1 end (a column-border intercept of P(ta)P(tb)):
// its slope dy/dx and the border's position v along it.
__device__ __forceinline__ float2 chord_terms(const EndSrc &e, const Trav &tv)
{
    if (e.kind != 1) return make_float2(0.0f, 0.0f);
    const float2 pa = bez(tv, e.ta), pb = bez(tv, e.tb);
    const float dxab = pb.x - pa.x;
    return make_float2((pb.y - pa.y) / dxab, (e.X - pa.x) / dxab);   // (sl, v)
}

template <int kMode>
__device__ void add_line_o(Queue &q, const Column &c, float2 l, float2 r, uint32_t bit, bool &added,
                           bool &enclosed, float4 *orig, float4 *ends, uint32_t &n_orig, const AdjDesc &d,
                           const EndSrc &el, const Trav &tl, const EndSrc &er, const Trav &tr)
{
    Thr t{cl_min(l.y, r.y), cl_max(l.y, r.y), l.x, r.x};
    const bool pos = l.y <= r.y;
    const bool persistent = (l.x != r.x) && (l.x == 0.0f);
    uint32_t h = (pos ? kPositiveSlope : 0u) | (persistent ? kPersist : 0u) | bit;
    const bool keep = hdr_persist(h) || ((t.bottom - t.top) >= kMinCrop);
    if (!keep) return;
    enclosed = enclosed || ((t.top <= 0.0f) && hdr_persist_top(h)) ||
               ((t.bottom <= 0.0f) && hdr_persist_bottom(h));
    if (!(t.top < c.float_height && t.bottom > 0.0f && t.left < 1.0f)) return;
    if (t.top <= 0.0f) {
        const float sx = intersect_x(h, t, 0.0f);
        if (hdr_pos(h)) { t = Thr{0.0f, t.bottom, sx, t.right}; h &= kUnpersist; }
        else { t = Thr{0.0f, t.bottom, t.left, sx}; }
    }
    if (t.right <= 0.0f) {
        enclosed = true;
        return;
    }
    added = true;
    const uint32_t k = n_orig++;
    if (kMode == kRecord) {
        if (q.push_slot()) {
            q.set(0, h | (k << kOriginShift), t);
            orig[2 * k * kWave] = make_float4(l.y, r.y, 0.0f, 0.0f);
            orig[(2 * k + 1) * kWave] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (ends) {
                const float kk = r.y != l.y ? (r.x - l.x) / (r.y - l.y) : 0.0f;
                const float2 cl = chord_terms(el, tl), cr = chord_terms(er, tr);
                ends[4 * k * kWave] = make_float4(kk, __int_as_float(el.curve), __int_as_float(el.kind), el.ta);
                ends[(4 * k + 1) * kWave] = make_float4(el.tb, cl.x, cl.y, __int_as_float(er.curve));
                ends[(4 * k + 2) * kWave] = make_float4(__int_as_float(er.kind), er.ta, er.tb, cr.x);
                ends[(4 * k + 3) * kWave] = make_float4(cr.y, 0.0f, 0.0f, 0.0f);
            }
        }
    } else {
        const float4 o = orig[2 * k * kWave];
        const float4 e = orig[(2 * k + 1) * kWave];
        if (o.z == 0.0f && o.w == 0.0f && e.x == 0.0f && e.y == 0.0f) return;
        const float gbx = o.w, gax = o.z - o.w;
        const float kk = r.y != l.y ? (r.x - l.x) / (r.y - l.y) : 0.0f;
        apply_end(d, el, tl, gax, -kk * gax + e.x);
        apply_end(d, er, tr, gbx, -kk * gbx + e.y);
    }
}

// spawnThresholds with end point origins.
template <int kMode>
__device__ void spawn_o(Queue &q, const Column &c, uint32_t bit, const Trav &l, const Trav &r, int lcur,
                        int rcur, bool &added, bool &enclosed, float4 *orig, float4 *ends, uint32_t &n_orig,
                        const AdjDesc &d)
{
    float yl;
    EndSrc e1{lcur, 0, 0.0f, 0.0f, 0.0f};
    if (l.left.x >= 0.0f) {
        yl = l.left.y;
    } else {
        const Cut k = intersect_track(l);
        yl = k.y;
        if (k.vertical) e1 = EndSrc{lcur, 2, k.min_at_b ? k.tb : k.ta, 0.0f, 0.0f};
        else e1 = EndSrc{lcur, 1, k.ta, k.tb, l.xpos};
    }
    const EndSrc e2{lcur, 0, 1.0f, 0.0f, 0.0f};
    const EndSrc e3{rcur, 0, 0.0f, 0.0f, 0.0f};
    float yr;
    EndSrc e4{rcur, 0, 1.0f, 0.0f, 0.0f};
    bool lw = false, rw = false;
    if ((l.right.x < 1.0f) && (l.right.x > 0.0f)) {
        add_line_o<kMode>(q, c, make_float2(l.xpos, yl), l.right, bit, added, enclosed, orig, ends, n_orig, d,
                          e1, l, e2, l);
        lw = true;
    }
    if (r.right.x <= 1.0f) {
        yr = r.right.y;
    } else {
        const Cut k = intersect_track(r);
        yr = k.y;
        if (k.vertical) e4 = EndSrc{rcur, 2, k.min_at_b ? k.tb : k.ta, 0.0f, 0.0f};
        else e4 = EndSrc{rcur, 1, k.ta, k.tb, r.xpos};
    }
    if ((r.left.x > 0.0f) && (r.left.x < 1.0f) && (l.index != r.index)) {
        add_line_o<kMode>(q, c, r.left, make_float2(r.xpos, yr), bit, added, enclosed, orig, ends, n_orig, d,
                          e3, r, e4, r);
        rw = true;
    }
    if (l.right.x < r.left.x || (!lw && !rw)) {
        const bool bl_right = lw || (l.left.x == l.right.x);
        const bool br_left = rw || (r.left.x == r.right.x);
        const float2 bl = bl_right ? l.right : make_float2(l.xpos, yl);
        const float2 br = br_left ? r.left : make_float2(r.xpos, yr);
        add_line_o<kMode>(q, c, bl, br, bit, added, enclosed, orig, ends, n_orig, d, bl_right ? e2 : e1, l,
                          br_left ? e3 : e4, r);
    }
}

// buildThresholdArray with origins: the same shape walk, prefilter and
// labels as the raster path's generate, so records and their order match it.
template <bool kCompact, int kMode>
__device__ __forceinline__ void generate_o(const Frame &f, const AdjDesc &d, Queue &q, const Column &c,
                                           bool active, int lane, float span0, float span1, int32_t *sidx,
                                           uint32_t &shape_bits, uint64_t &stack0, uint64_t &stack1,
                                           float4 *orig, float4 *ends = nullptr, uint32_t *lines = nullptr)
{
    const SceneDesc &s = d.s;
    const float dx = (float)c.dx, dy = (float)c.dy;
    const uint32_t max_shapes = (uint32_t)f.max_shapes;
    const int num_shapes = (int)__builtin_amdgcn_readfirstlane(c.num_shapes);
    const int shape_start = (int)__builtin_amdgcn_readfirstlane(c.shape_start);
    const float below = dy + c.float_height + 1.0f;
    uint32_t n_orig = 0;
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
                const GUDNI_CONST uint32_t *hw = reinterpret_cast<const GUDNI_CONST uint32_t *>(strand);
                const uint32_t size = hw[0] & 0xFFFFu;
                const int curve_base = (int)hw[1];
                bool enclosed_strand = false;
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
                    int lcur, rcur;
                    search_curve_pair(l, r, strand + 32, tree_size, dx, dy, lcur, rcur);
                    lcur += curve_base;
                    rcur += curve_base;
                    spawn_o<kMode>(q, c, label, l, r, lcur, rcur, added, enclosed_strand, orig, ends, n_orig,
                                   d);
                }
                offset += 8u * size;
                enclosed_shape = enclosed_shape != enclosed_strand;
            }
            if (enclosed_shape) {
                if (label < 64u) stack0 ^= 1ull << label;
                else stack1 ^= 1ull << (label - 64u);
            }
            if (added || enclosed_shape) {
                if (!kCompact && shape_bits < max_shapes) sidx[(int)shape_bits * kWave] = shape_index;
                shape_bits += 1u;
            }
        }
    }
    if (lines) *lines = n_orig;
}

// Gradient (gx, gy) on a recorded line end -> outline points (apply_end with
// the chord terms recorded at generation).
__device__ __forceinline__ void apply_recorded_end(const AdjDesc &d, int curve, int kind, float ta, float tb,
                                                   float sl, float v, float gx, float gy)
{
    if (kind == 0) {
        chain_point(d, curve, ta, make_float2(gx, gy));
    } else if (kind == 2) {
        chain_point(d, curve, ta, make_float2(0.0f, gy));
    } else {
        chain_point(d, curve, ta, make_float2(-gy * sl * (1.0f - v), gy * (1.0f - v)));
        chain_point(d, curve, tb, make_float2(-gy * sl * v, gy * v));
    }
}

// Phase C from the recorded ends: every generated line's accumulated
// weights -> its end points -> outline points, in generation order, as the
// regenerating walk (generate_o<kApply>) applies them.
__device__ __forceinline__ void apply_lines(const AdjDesc &d, const float4 *orig, const float4 *ends,
                                            uint32_t n_lines)
{
    for (uint32_t k = 0; k < n_lines; ++k) {
        const float4 o = orig[2 * k * kWave];
        const float4 e = orig[(2 * k + 1) * kWave];
        if (o.z == 0.0f && o.w == 0.0f && e.x == 0.0f && e.y == 0.0f) continue;
        const float4 r0 = ends[4 * k * kWave], r1 = ends[(4 * k + 1) * kWave];
        const float4 r2 = ends[(4 * k + 2) * kWave], r3 = ends[(4 * k + 3) * kWave];
        const float kk = r0.x;
        const float gbx = o.w, gax = o.z - o.w;
        apply_recorded_end(d, __float_as_int(r0.y), __float_as_int(r0.z), r0.w, r1.x, r1.y, r1.z, gax,
                           -kk * gax + e.x);
        apply_recorded_end(d, __float_as_int(r1.w), __float_as_int(r2.x), r2.y, r2.z, r2.w, r3.x, gbx,
                           -kk * gbx + e.y);
    }
}

// Straight (non-premultiplied) colour at pixel (ax, ay) of the shape
// labelled b, and its tag; a compact tile reads the substance record from
// the tile's LDS copy.
template <bool kCompact>
__device__ __forceinline__ float4 label_color(const SceneDesc &s, const int32_t *sidx, const TileTable *tt,
                                              const float4 *straight, int b, int ax, int ay, uint64_t &tag,
                                              LayerCache &lc)
{
    if (kCompact) {
        tag = tt->tag[b];
        if ((tag & kTagTypeMask) == kTagSolid) return straight[b];
        return lc.color(s, tt, b, tag, straight[b], ax, ay);
    }
    tag = s.shapes[sidx[b * kWave]].tag;
    return read_color(s, tag, ax, ay);
}

// determineColor's merge of consecutive shapes with one substance id
// (Kernels.cl:1447-1513), as a walk top-down: visit(tag) says whether the
// layer is composited.  A new id composites if its shape adds (or
// continues); a repeated id never composites again, and only the deprecated
// continue compound toggles whether the merged run is set.
struct MergeWalk {
    uint64_t last_id = ~0ull;
    bool last_is_continue = true, last_is_set = false;
    __device__ __forceinline__ bool visit(uint64_t tag)
    {
        const uint64_t sid = tag & kTagIdMask;
        const bool is_cont = (tag & kTagCompoundMask) == kTagContinue;
        const bool is_add = (tag & kTagCompoundMask) == kTagAdd;
        if (sid == last_id) {
            if (last_is_continue) {
                if (!is_cont) { last_is_set = is_add; last_is_continue = false; }
                else last_is_set = !last_is_set;
            }
            return false;
        }
        last_id = sid;
        last_is_set = is_add || is_cont;
        return last_is_set;
    }
};

__device__ __forceinline__ uint64_t label_tag(const SceneDesc &s, const int32_t *sidx, const TileTable *tt,
                                              bool compact, int b)
{
    return compact ? tt->tag[b] : s.shapes[sidx[b * kWave]].tag;
}

// Colour gradient of one section: gc = dL/dC for C = (P/A, A), S = (P, A)
// the section's premultiplied composite.  Walks the stack top-down as the
// forward does (stopping where its alpha reached 1).  For layer i with
// transmittance T and straight colour (c, a): dS/dc = T a, dS/da = T (c, 1) -
// V with V = T * (composite below i) = (S - F_{i+1}) / (1 - a), F_{i+1} the
// composite down to and including i; an opaque layer takes V from a walk of
// the layers below it.  Layers the merge walk does not composite (a repeated
// substance id, a subtract compound) add no term.
template <bool kCompact>
__device__ void color_grad(const SceneDesc &s, const Stack &st, const int32_t *sidx, const TileTable *tt,
                           const float4 *straight, float4 *dcol, float *d_rgba, int ax, int ay, float4 S,
                           float4 C, float4 gc, LayerCache &lc)
{
    if (!(S.w > 0.0f)) return;
    const float gPx = gc.x / S.w, gPy = gc.y / S.w, gPz = gc.z / S.w;
    const float gA = gc.w - (gc.x * C.x + gc.y * C.y + gc.z * C.z) / S.w;
    uint64_t w0 = st.s0, w1 = st.s1;
    float4 F = make_float4(0.f, 0.f, 0.f, 0.f);
    MergeWalk mw;
    for (;;) {
        int b;
        if (w1) { b = 127 - __clzll((long long)w1); w1 &= ~(1ull << (b - 64)); }
        else if (w0) { b = 63 - __clzll((long long)w0); w0 &= ~(1ull << b); }
        else break;
        if (!mw.visit(label_tag(s, sidx, tt, kCompact, b))) continue;
        uint64_t tag;
        const float4 col = label_color<kCompact>(s, sidx, tt, straight, b, ax, ay, tag, lc);
        const float T = 1.0f - F.w;
        const float4 pm = make_float4(col.x * col.w, col.y * col.w, col.z * col.w, col.w);
        float4 Fn;
        Fn.x = F.x + pm.x * T;
        Fn.y = F.y + pm.y * T;
        Fn.z = F.z + pm.z * T;
        Fn.w = F.w + pm.w * T;
        if ((tag & kTagTypeMask) == kTagSolid && T > 0.0f) {
            float4 V;
            if (col.w < 1.0f) {
                const float inv = 1.0f / (1.0f - col.w);
                V = make_float4((S.x - Fn.x) * inv, (S.y - Fn.y) * inv, (S.z - Fn.z) * inv, (S.w - Fn.w) * inv);
            } else {
                // composite of the layers below, scaled by T (the merge
                // walk goes on from this layer)
                uint64_t u0 = w0, u1 = w1;
                float4 U = make_float4(0.f, 0.f, 0.f, 0.f);
                bool reached_one = false;
                MergeWalk mu = mw;
                while (u0 | u1) {
                    int bb;
                    if (u1) { bb = 127 - __clzll((long long)u1); u1 &= ~(1ull << (bb - 64)); }
                    else { bb = 63 - __clzll((long long)u0); u0 &= ~(1ull << bb); }
                    if (!mu.visit(label_tag(s, sidx, tt, kCompact, bb))) continue;
                    uint64_t tag2;
                    const float4 c2 = label_color<kCompact>(s, sidx, tt, straight, bb, ax, ay, tag2, lc);
                    const float t2 = 1.0f - U.w;
                    U.x = U.x + c2.x * c2.w * t2;
                    U.y = U.y + c2.y * c2.w * t2;
                    U.z = U.z + c2.z * c2.w * t2;
                    U.w = U.w + c2.w * t2;
                    if (U.w == 1.0f) { reached_one = true; break; }
                }
                if (!reached_one) {
                    const float t2 = 1.0f - U.w;
                    U.x = U.x + s.bg[0] * s.bg[3] * t2;
                    U.y = U.y + s.bg[1] * s.bg[3] * t2;
                    U.z = U.z + s.bg[2] * s.bg[3] * t2;
                    U.w = U.w + s.bg[3] * t2;
                }
                V = make_float4(T * U.x, T * U.y, T * U.z, T * U.w);
            }
            const float ta = T * col.w;
            const float gr = gPx * ta, gg = gPy * ta, gb = gPz * ta;
            const float ga = T * (gPx * col.x + gPy * col.y + gPz * col.z + gA) -
                             (gPx * V.x + gPy * V.y + gPz * V.z + gA * V.w);
            if (kCompact) {
                float *p = reinterpret_cast<float *>(dcol + b);
                atomicAdd(p + 0, gr);
                atomicAdd(p + 1, gg);
                atomicAdd(p + 2, gb);
                atomicAdd(p + 3, ga);
            } else {
                float *p = d_rgba + 4 * (size_t)(tag & kTagIdMask);
                atomicAdd(p + 0, gr);
                atomicAdd(p + 1, gg);
                atomicAdd(p + 2, gb);
                atomicAdd(p + 3, ga);
            }
        }
        F = Fn;
        if (F.w == 1.0f) break;
    }
}

// Mid-height x of threshold `cur` in the current span, before clamping
// (horizontalAdvance / thresholdMidXLow in accurate arithmetic).
__device__ __forceinline__ float mid_x(const Queue &q, const Sweep &p, int cur)
{
    const Thr t = q.get(cur);
    const uint32_t h = q.get_h(cur);
    if (t.left == t.right) return t.left;
    const float sy = hdr_pos(h) ? t.top : t.bottom;
    const float dd = (p.sy0 - sy) + ((p.sy1 - p.sy0) * 0.5f);
    return (dd * inv_slope(h, t)) + t.left;
}

// The generated line end point a span boundary sits on: an active record
// ending there or a waiting record starting there whose line has an end
// point at exactly that height.
__device__ void boundary_origin(const Queue &q, const Sweep &p, const float4 *orig, int &org, int &end)
{
    const float y = p.sy1;
    for (int i = 0; i < q.len; ++i) {
        const Thr t = q.get(i);
        if (i >= p.num_active && t.top > y) break;
        if ((i < p.num_active ? t.bottom : t.top) != y) continue;
        const uint32_t o = (q.get_h(i) >> kOriginShift) & kOriginMask;
        const float4 ov = orig[2 * o * kWave];
        if (ov.x == y) { org = (int)o; end = 0; return; }
        if (ov.y == y) { org = (int)o; end = 1; return; }
    }
}

// dL/dy of the end point a span boundary sits on.
__device__ __forceinline__ void span_term(float4 *orig, int org, int end, const float4 G, const float4 above,
                                          const float4 below)
{
    const float J = G.x * (above.x - below.x) + G.y * (above.y - below.y) + G.z * (above.z - below.z) +
                    G.w * (above.w - below.w);
    if (J == 0.0f) return;
    float4 *e = orig + (2 * org + 1) * kWave;
    if (end == 0) e->x += J;
    else e->y += J;
}

// Forward sweep: the image, per-pixel dL/d(accumulated colour), the loss.
template <bool kCompact>
__device__ __forceinline__ uint32_t sweep_forward(const Frame &f, const AdjDesc &d, Queue &q, const Column &c,
                                                  Stack st, const int32_t *sidx, const TileTable *tt, float &loss)
{
    const SceneDesc &s = d.s;
    Sweep p;
    p.current = 0; p.num_active = 0;
    p.sx0 = 0.0f; p.sy0 = 0.0f; p.sx1 = 1.0f; p.sy1 = 0.0f;
    p.rnd = 0;
    long guard = kSectionGuard;
    int ay = c.dy;
    // picture and gradient colours of the current pixel (compact tiles)
    LayerCache lc{const_cast<int32_t *>(sidx), 0u};
    // one loop over sections and pixel ends, as render's
    p.pixel_y = 1.0f;
    if (!(p.pixel_y <= c.float_height)) return 0;
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float area_sum = 0.0f;
    for (;;) {
        if ((p.sx1 < 1.0f) || (p.sy1 < p.pixel_y)) {
            if (--guard < 0) return kStatusGuard;
            vertical_advance(q, c, p, st);
            if (q.spilled || q.overflow) return 0;
            float nx = 1.0f;
            if (p.current < p.num_active) {
                const float x = mid_x(q, p, p.current);
                nx = (x >= 1.0f) ? 0.0f : cl_max(0.0f, x);
            }
            p.sx0 = p.sx1;
            p.sx1 = nx;
            bool positional = false;
            const float4 col = determine_color<kCompact>(f, s, st, sidx, tt, c.dx, ay, positional,
                                                         kCompact ? &lc : nullptr);
            const float area = (p.sx1 - p.sx0) * (p.sy1 - p.sy0);
            acc.x += col.x * area;
            acc.y += col.y * area;
            acc.z += col.z * area;
            acc.w += col.w * area;
            area_sum += area;
            if (p.current < p.num_active) st.flip(q.get_h(p.current));
            ++p.current;
        } else {
            const float4 color = make_float4(acc.x / area_sum, acc.y / area_sum, acc.z / area_sum,
                                             acc.w / area_sum);
            const long long pix = (long long)ay * s.width + c.dx;
            s.out_f32[pix] = color;
            const float4 tg = d.target[pix];
            const float4 e = make_float4(color.x - tg.x, color.y - tg.y, color.z - tg.z, color.w - tg.w);
            loss += e.x * e.x + e.y * e.y + e.z * e.z + e.w * e.w;
            const float k = 2.0f / area_sum;
            d.gpix[pix] = make_float4(k * e.x, k * e.y, k * e.z, k * e.w);
            p.sx0 = 0.0f;
            p.sy0 = p.pixel_y;
            ++ay;
            p.pixel_y += 1.0f;
            if (!(p.pixel_y <= c.float_height)) break;
            acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            area_sum = 0.0f;
            lc.known = 0u;
        }
    }
    return 0;
}

// Backward sweep over the same sections.
template <bool kCompact>
__device__ __forceinline__ uint32_t sweep_backward(const Frame &f, const AdjDesc &d, Queue &q, const Column &c,
                                                   Stack st, const int32_t *sidx, const TileTable *tt,
                                                   const float4 *straight, float4 *dcol, float4 *orig)
{
    const SceneDesc &s = d.s;
    Sweep p;
    p.current = 0; p.num_active = 0;
    p.sx0 = 0.0f; p.sy0 = 0.0f; p.sx1 = 1.0f; p.sy1 = 0.0f;
    p.rnd = 0;
    long guard = kSectionGuard;
    int ay = c.dy;
    LayerCache lc{const_cast<int32_t *>(sidx), 0u};
    // one loop over sections and pixel ends, as render's; the span state
    // below is reset at each pixel's start
    p.pixel_y = 1.0f;
    if (!(p.pixel_y <= c.float_height)) return 0;
    float4 G = d.gpix[(long long)ay * s.width + c.dx];
    bool have_prev = false, prev_inside = false;
    uint32_t prev_origin = 0;
    float4 prev_col = make_float4(0.f, 0.f, 0.f, 0.f);
    // span boundaries of this row: sum C (e_j - e_{j-1}) along the span's
    // top and bottom edges; `pend` is the boundary at the current span's
    // top with the span above's bottom sum and the end point it moves with
    float4 top_sum = make_float4(0.f, 0.f, 0.f, 0.f), bot_sum = top_sum, pend_bot = top_sum;
    float e_top = 0.0f, e_bot = 0.0f;
    bool span_open = false, pend = false;
    int pend_org = -1, pend_end = 0, cur_org = -1, cur_end = 0;
    for (;;) {
        if ((p.sx1 < 1.0f) || (p.sy1 < p.pixel_y)) {
            if (--guard < 0) return kStatusGuard;
            vertical_advance(q, c, p, st);
            if (q.spilled || q.overflow) return 0;
            if (p.current == 0) {   // a new span starts at the column's left edge
                have_prev = false;
                if (span_open && pend && pend_org >= 0)
                    span_term(orig, pend_org, pend_end, G, pend_bot, top_sum);
                if (span_open) {
                    pend = true;
                    pend_org = cur_org;
                    pend_end = cur_end;
                    pend_bot = bot_sum;
                }
                span_open = true;
                top_sum = make_float4(0.f, 0.f, 0.f, 0.f);
                bot_sum = top_sum;
                e_top = e_bot = 0.0f;
                cur_org = -1;
                if (p.sy1 != p.pixel_y && p.sy1 != c.float_height)
                    boundary_origin(q, p, orig, cur_org, cur_end);
            }
            float nx = 1.0f, x = 1.0f;
            if (p.current < p.num_active) {
                x = mid_x(q, p, p.current);
                nx = (x >= 1.0f) ? 0.0f : cl_max(0.0f, x);
            }
            p.sx0 = p.sx1;
            p.sx1 = nx;
            bool positional = false;
            const float4 col = determine_color<kCompact>(f, s, st, sidx, tt, c.dx, ay, positional,
                                                         kCompact ? &lc : nullptr);
            const float dy = p.sy1 - p.sy0;
            const float area = (p.sx1 - p.sx0) * dy;
            {
                float et = 1.0f, eb = 1.0f;
                if (p.current < p.num_active) {
                    if (x > 0.0f && x < 1.0f) {
                        const Thr t = q.get(p.current);
                        const uint32_t h = q.get_h(p.current);
                        et = intersect_x(h, t, p.sy0);
                        eb = intersect_x(h, t, p.sy1);
                    } else {
                        et = eb = nx;
                    }
                }
                top_sum.x += col.x * (et - e_top); top_sum.y += col.y * (et - e_top);
                top_sum.z += col.z * (et - e_top); top_sum.w += col.w * (et - e_top);
                bot_sum.x += col.x * (eb - e_bot); bot_sum.y += col.y * (eb - e_bot);
                bot_sum.z += col.z * (eb - e_bot); bot_sum.w += col.w * (eb - e_bot);
                e_top = et;
                e_bot = eb;
            }
            if (have_prev && prev_inside && dy > 0.0f) {
                const float w = (G.x * (prev_col.x - col.x) + G.y * (prev_col.y - col.y) +
                                 G.z * (prev_col.z - col.z) + G.w * (prev_col.w - col.w)) * dy;
                if (w != 0.0f) {
                    float4 *o = orig + 2 * prev_origin * kWave;
                    const float4 ov = *o;
                    const float ym = p.sy0 + dy * 0.5f;
                    const float u = ov.y != ov.x ? (ym - ov.x) / (ov.y - ov.x) : 0.0f;
                    o->z = ov.z + w;
                    o->w = ov.w + w * u;
                }
            }
            const float4 gc = make_float4(G.x * area, G.y * area, G.z * area, G.w * area);
            if (gc.x != 0.0f || gc.y != 0.0f || gc.z != 0.0f || gc.w != 0.0f) {
                const float4 S = make_float4(col.x * col.w, col.y * col.w, col.z * col.w, col.w);
                color_grad<kCompact>(s, st, sidx, tt, straight, dcol, d.d_rgba, c.dx, ay, S, col, gc, lc);
            }
            if (p.current < p.num_active) {
                const uint32_t h = q.get_h(p.current);
                prev_origin = (h >> kOriginShift) & kOriginMask;
                prev_inside = x > 0.0f && x < 1.0f;
                prev_col = col;
                have_prev = true;
                st.flip(h);
            } else {
                have_prev = false;
            }
            ++p.current;
        } else {
            if (span_open && pend && pend_org >= 0) span_term(orig, pend_org, pend_end, G, pend_bot, top_sum);
            p.sx0 = 0.0f;
            p.sy0 = p.pixel_y;
            ++ay;
            p.pixel_y += 1.0f;
            if (!(p.pixel_y <= c.float_height)) break;
            G = d.gpix[(long long)ay * s.width + c.dx];
            lc.known = 0u;
            have_prev = false;
            prev_inside = false;
            prev_origin = 0;
            prev_col = make_float4(0.f, 0.f, 0.f, 0.f);
            top_sum = make_float4(0.f, 0.f, 0.f, 0.f);
            bot_sum = top_sum;
            pend_bot = top_sum;
            e_top = e_bot = 0.0f;
            span_open = false;
            pend = false;
            pend_org = -1;
            pend_end = 0;
            cur_org = -1;
            cur_end = 0;
        }
    }
    return 0;
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One work unit, phases A-D.  It and the stages it calls (generation, both
// sweeps, the line map) are inlined into the kernel: a call would take the
// frame, the scene descriptor and the queues by address, which keeps them in
// per-lane scratch memory for the whole unit (848 bytes a lane with 46
// spilled VGPRs, against 80 bytes and none inlined).
template <bool kCompact>
__device__ __forceinline__ void adj_unit(const Frame &f, const AdjDesc &d, int lane, uint32_t unit, Queue &q,
                                         Queue &fq, int32_t *sidx, TileTable *tt, float4 *straight,
                                         float4 *dcol, float4 *orig, float4 *ends, uint32_t &status,
                                         float &loss)
{
    const SceneDesc &s = d.s;
    const int tile = (int)(unit >> 2), quarter = (int)(unit & 3u);
    const TileRec t = f.tiles[tile];
    const int job = f.tile_job ? f.tile_job[tile] : 0;
    const int column = unit_column(t.hdepth, f.compute_depth, quarter, lane);
    Column c;
    const bool active = init_column(f, s, t, job, column, c);
    if (kCompact) {
        for (int n = lane; n < c.num_shapes; n += kWave) {
            const uint64_t tag = s.shapes[c.shape_start + n].tag;
            tt->tag[n] = tag;
            float4 col = s.substances[tag & kTagIdMask];
            straight[n] = col;
            if ((tag & kTagTypeMask) == kTagSolid)
                col = make_float4(col.x * col.w, col.y * col.w, col.z * col.w, col.w);
            tt->color[n] = col;
            dcol[n] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        // the tile-local shapes whose substance is a picture or gradient
        uint64_t pic_lo = 0, pic_hi = 0;
        for (int n = lane; n < c.num_shapes; n += kWave) {
            if ((s.shapes[c.shape_start + n].tag & kTagTypeMask) != kTagSolid) {
                if (n < 64) pic_lo |= 1ull << n; else pic_hi |= 1ull << (n - 64);
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            pic_lo |= __shfl_xor(pic_lo, o);
            pic_hi |= __shfl_xor(pic_hi, o);
        }
        if (lane == 0) {
            tt->pic0 = pic_lo;
            tt->pic1 = pic_hi;
        }
        wave_sync();
    }
    const int x0 = t.box[0] + unit_x0(t.hdepth, f.compute_depth, quarter);
    const int x1 = x0 + unit_width(t.hdepth, f.compute_depth);
    const float span0 = (float)x0 - 1.0f, span1 = (float)x1 + 1.0f;
    // optional phase clocks (wave-uniform points): [0] A generate [1] A sort
    // [2] A sweep [3] B generate [4] B sort [5] B sweep [6] C [7] units
    unsigned long long *clk = f.stage_cycles;
    uint16_t *hist = reinterpret_cast<uint16_t *>(&tt->gc[0][0][0]) + lane;   // sort histogram
    long long tclk = clk ? clock64() : 0;
    auto stamp = [&](int i) {
        if (!clk) return;
        const long long now = clock64();
        if (lane == 0) atomicAdd(clk + i, (unsigned long long)(now - tclk));
        tclk = now;
    };
    // A: forward
    // every generate_o call is made by the whole wave (it ballots over shapes)
    uint32_t bits = 0, n_lines = 0;
    uint64_t s0 = 0, s1 = 0;
    generate_o<kCompact, kRecord>(f, d, q, c, active, lane, span0, span1, sidx, bits, s0, s1, orig, ends,
                                  &n_lines);
    stamp(0);
    bool ok = active && !q.spilled && !q.overflow;
    if (ok) radix_sort_queue(q, hist, s.sort_bits, c.float_height);
    stamp(1);
    // the forward sweep splits and pops records, and runs on fq: with a ring
    // of its own, fq holds a copy of the sorted queue and q stays as the
    // backward sweep needs it (the same records a regeneration would give,
    // at the cost of one copy); without one, fq is q's ring and the backward
    // sweep regenerates.  (One queue object per role keeps both in
    // registers.)  A forward sweep that outgrows its ring hands the unit to
    // the retry pass like a full generation: its flag goes to q.
    const bool copied = fq.thr != q.thr;
    fq.spilled = false;
    fq.overflow = false;
    if (copied) {
        fq.head = 0;
        fq.len = 0;
        if (ok) {
            for (int i = 0; i < q.len; ++i) fq.set(i, q.get_h(i), q.get(i));
            fq.len = q.len;
        }
    } else {
        fq.head = q.head;
        fq.len = q.len;
    }
    if (ok) status = sweep_forward<kCompact>(f, d, fq, c, Stack{s0, s1, 0u}, sidx, tt, loss);
    stamp(2);
    q.spilled = q.spilled || fq.spilled;
    if (active && (q.overflow || fq.overflow)) status |= kStatusOverflow;
    if (__any(q.spilled) || __any(status != 0)) return;
    // B: the sorted records again (kept, or recomputed), backward sweep
    if (!copied) {
        q.head = 0; q.len = 0;
        bits = 0; s0 = 0; s1 = 0;
        generate_o<kCompact, kRecord>(f, d, q, c, active, lane, span0, span1, sidx, bits, s0, s1, orig, ends,
                                      &n_lines);
    }
    stamp(3);
    ok = active && !q.spilled && !q.overflow;
    if (ok && !copied) radix_sort_queue(q, hist, s.sort_bits, c.float_height);
    stamp(4);
    if (ok) status = sweep_backward<kCompact>(f, d, q, c, Stack{s0, s1, 0u}, sidx, tt, straight, dcol, orig);
    stamp(5);
    // B repeats A's queue operations on the same records, so it cannot
    // outgrow a ring the forward sweep fit in; should it, the unit goes to
    // the retry pass before C and D flush anything, so that nothing is
    // counted twice
    if (__any(q.spilled)) return;
    // C: line weights -> outline points, from the ends recorded in A (or
    // by walking the generation once more)
    if (ends) {
        if (ok && status == 0) apply_lines(d, orig, ends, n_lines);
    } else {
        uint32_t bits2 = 0;
        uint64_t a0 = 0, a1 = 0;
        generate_o<kCompact, kApply>(f, d, q, c, ok && status == 0, lane, span0, span1, sidx, bits2, a0,
                                     a1, orig);
    }
    stamp(6);
    if (clk && lane == 0) atomicAdd(clk + 7, 1ull);
    // D: the tile's colour gradients -> substances
    if (kCompact) {
        wave_sync();
        for (int n = lane; n < c.num_shapes; n += kWave) {
            const uint64_t tag = tt->tag[n];
            if ((tag & kTagTypeMask) != kTagSolid) continue;
            const float4 g = dcol[n];
            float *p = d.d_rgba + 4 * (size_t)(tag & kTagIdMask);
            if (g.x != 0.0f) atomicAdd(p + 0, g.x);
            if (g.y != 0.0f) atomicAdd(p + 1, g.y);
            if (g.z != 0.0f) atomicAdd(p + 2, g.z);
            if (g.w != 0.0f) atomicAdd(p + 3, g.w);
        }
        wave_sync();
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) adjoint_pass(AdjFrame a, int from_retry)
{
    __shared__ TileTable tables[4];
    __shared__ float4 dcols[4][128];
    __shared__ float4 straights[4][128];   // compact tiles: the shapes' substance records
    const Frame &f = a.f;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int slot = blockIdx.x * (blockDim.x / kWave) + wave;
    const int cap = 1 << f.cap_log2;
    float4 *q_thr = f.q_thr + ((size_t)slot * kWave + lane) * cap;
    uint32_t *q_hdr = f.q_hdr + ((size_t)slot * kWave + lane) * cap;
    float4 *orig = a.q_orig + (size_t)slot * 2 * cap * kWave + lane;
    float4 *ends = a.q_ends ? a.q_ends + (size_t)slot * 4 * cap * kWave + lane : nullptr;
    float4 *c_thr = a.copy_thr ? a.copy_thr + ((size_t)slot * kWave + lane) * cap : q_thr;
    uint32_t *c_hdr = a.copy_hdr ? a.copy_hdr + ((size_t)slot * kWave + lane) * cap : q_hdr;
    int32_t *sidx = f.shape_idx + (size_t)slot * kSidxWords * kWave + lane;
    TileTable *tt = &tables[wave];
    float4 *dcol = dcols[wave];
    float4 *straight = straights[wave];
    const int n_work = from_retry ? (int)min((uint32_t)f.retry_capacity, f.retry[0]) : f.n_work;
    Claimer claim;
    claim.init(f.counter, f.n_parts, n_work, (int)blockIdx.x);
    for (;;) {
        const int w = claim.next(lane);
        if (w < 0) break;
        const uint32_t unit = from_retry ? f.retry[1 + w] : f.work[w];
        const int count = f.tiles[unit >> 2].shape_count;
        const int scene = f.tile_scene
                              ? (int)__builtin_amdgcn_readfirstlane((uint32_t)f.tile_scene[unit >> 2])
                              : 0;
        // wave-uniform, read through the scalar cache where a field is used
        // (a register copy spilled 68 more scalar registers; the 1024^2 step:
        // 6.60 -> 6.37 ms)
        const AdjDesc &d = *(const AdjDesc *)ro(a.scenes + scene);
        Queue q;
        q.thr = q_thr; q.hdr = q_hdr;
        q.head = 0; q.len = 0; q.mask = cap - 1; q.maxt = f.max_thresholds;
        q.spilled = false; q.overflow = false;
        uint32_t status = 0;
        float loss = 0.0f;
        Queue fq = q;
        fq.thr = c_thr; fq.hdr = c_hdr;
        if (count <= f.max_shapes)
            adj_unit<true>(f, d, lane, unit, q, fq, sidx, tt, straight, dcol, orig, ends, status, loss);
        else
            adj_unit<false>(f, d, lane, unit, q, fq, sidx, tt, straight, dcol, orig, ends, status, loss);
        if (status) atomicOr(f.status, status);
        const bool spilled = __any(q.spilled);
        if (spilled) {
            if (lane == 0 && !from_retry) {
                const uint32_t k = atomicAdd(f.retry, 1u);
                if ((int)k < f.retry_capacity) f.retry[1 + k] = unit;
                else atomicOr(f.status, kStatusOverflow);
            }
        } else {
            for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
            if (lane == 0 && loss != 0.0f) atomicAdd(d.loss, (double)loss);
        }
    }
}

__global__ void adj_clear(float *p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

// Scene-local work units (tile << 2 | quarter) moved to the batch's tile
// numbering, and the batch's per-tile scene index.
__global__ void adj_rebase_work(const uint32_t *src, uint32_t *dst, int n, uint32_t tile_base)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[i] = src[i] + (tile_base << 2);
}

__global__ void adj_fill_i32(int32_t *dst, int n, int32_t v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = v;
}

__global__ void adj_reset(uint32_t *counters, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counters[i] = 0u;
}

}  // namespace adj
}  // namespace gudni

// ===========================================================================
// Host side
// ===========================================================================
using namespace gudni;

struct gudni_adjoint_batch {
    struct Scene {
        int32_t width = 0, height = 0;
        float bg[4] = {1, 1, 1, 1};
        std::vector<uint64_t> tags;
        std::vector<int32_t> shape_off, outline_off;
        std::vector<float> pairs;
        std::vector<float4> subs;
        std::vector<uint8_t> pict, uses;
        std::vector<gudni_gradient> grads;
    };
    // A scene rebuilt on the device after a parameter update: its outline
    // structure (uploaded once), serialised geometry with curve provenance,
    // and its tile tree (scene-local work units).
    struct DevScene {
        DevBuf tags, soff, ooff;
        DevBuf geo, shapes, bbox, sbounds, curves;
        DevBuf tiles, tile_job, refs, ref_bounds, work;
        int32_t n_tiles = 0, n_work = 0, n_kept = 0;
        void release()
        {
            DevBuf *bufs[] = {&tags, &soff, &ooff, &geo, &shapes, &bbox, &sbounds, &curves,
                              &tiles, &tile_job, &refs, &ref_bounds, &work};
            for (DevBuf *b : bufs) b->release();
        }
    };
    void drop_device_scenes()
    {
        for (DevScene &v : dev) v.release();
        dev.clear();
    }
    int device = 0;
    std::vector<Scene> scenes;
    DevBuf geo, subs, pict, uses, shapes, bounds, curves, points, target, image, gpix, d_points, d_rgba,
        loss, tiles, tile_job, tile_scene, work, descs, orig_fast, orig_slow, grads, copy_thr, copy_hdr,
        ends_fast;
    std::vector<DevScene> dev;              // empty until the first device rebuild
    std::vector<size_t> geo_off, refs_off, curves_off, pict_off, uses_off, grads_off;
    std::vector<size_t> points_off, subs_off, pix_off;
    std::vector<int32_t> work_begin, work_count;
    int32_t n_work = 0;
    size_t n_points = 0, n_subs = 0, n_pix = 0;
    int blocks = 0;
    bool broken = false;     // a device update failed part-way
};

namespace {

// Requirements of the gradient walk: every shape is added (no subtract /
// continue compounds) and has a substance of its own.
int check_scene(gudni_ctx *ctx, const gudni_adjoint_batch::Scene &sc, int s)
{
    for (uint64_t tag : sc.tags) {
        if (tag & dev::kTagNonzero)
            return set_error(ctx, GUDNI_ERR_ARG, "adjoint scene " + std::to_string(s) +
                                                     ": the adjoint fills by crossing parity only "
                                                     "(no GUDNI_TAG_NONZERO shapes)");
    }
    return GUDNI_OK;
}

int fill_descs(gudni_ctx *ctx, gudni_adjoint_batch *a);

// Serialises, bins and validates every scene, then (re)uploads all buffers
// but the targets.
int build(gudni_ctx *ctx, gudni_adjoint_batch *a)
{
    std::vector<uint8_t> geo, pict, uses;
    std::vector<float4> subs;
    std::vector<ShapeRec> refs;
    std::vector<float> bounds, points;
    std::vector<CurveSrc> curves;
    std::vector<TileRec> tiles;
    std::vector<int32_t> tile_job, tile_scene;
    std::vector<uint32_t> work;
    std::vector<size_t> &geo_off = a->geo_off, &pict_off = a->pict_off, &uses_off = a->uses_off,
                        &refs_off = a->refs_off, &curves_off = a->curves_off, &grads_off = a->grads_off;
    geo_off.clear(); pict_off.clear(); uses_off.clear(); refs_off.clear(); curves_off.clear(); grads_off.clear();
    std::vector<gudni_gradient> grads;
    a->drop_device_scenes();
    a->points_off.clear(); a->subs_off.clear(); a->pix_off.clear();
    a->work_begin.clear(); a->work_count.clear();
    size_t pix = 0;
    const int n = (int)a->scenes.size();
    for (int s = 0; s < n; ++s) {
        const gudni_adjoint_batch::Scene &sc = a->scenes[s];
        gudni_outlines o;
        o.n_shapes = (int32_t)sc.tags.size();
        o.tags = sc.tags.data();
        o.shape_outline_offsets = sc.shape_off.data();
        o.outline_pair_offsets = sc.outline_off.data();
        o.pairs = sc.pairs.data();
        Geometry g;
        int rc = build_geometry(&o, sc.width, sc.height, g, true);
        if (rc) return set_error(ctx, rc, g_host_error);
        Jobs j;
        rc = build_jobs(ctx->spec, g, sc.width, sc.height, j);
        if (rc) return set_error(ctx, rc, g_host_error);
        const int n_jobs = (int)j.job_tile_offset.size() - 1;
        std::vector<int32_t> tj(j.tiles.size());
        for (int jb = 0; jb < n_jobs; ++jb)
            for (int i = j.job_tile_offset[jb]; i < j.job_tile_offset[jb + 1]; ++i) {
                j.tiles[i].shape_start += j.job_shape_offset[jb];
                tj[i] = n_jobs - 1 - jb;
            }
        const CommonView cv{g.heap.data(), g.heap.size(), sc.subs.data(), (int32_t)sc.subs.size(),
                            (int32_t)(sc.uses.size() / 24), (int32_t)sc.grads.size()};
        std::vector<float> xr(4 * j.refs.size() + 4);
        rc = validate_job(ctx, cv, j.refs.data(), (int32_t)j.refs.size(), j.tiles.data(),
                          (int32_t)j.tiles.size(), sc.width, sc.height, xr);
        if (rc) return set_error(ctx, rc, "adjoint scene " + std::to_string(s) + ": " + ctx->error);
        geo_off.push_back(geo.size());
        geo.insert(geo.end(), g.heap.begin(), g.heap.end());
        geo.resize((geo.size() + 15) & ~(size_t)15);
        a->subs_off.push_back(subs.size());
        subs.insert(subs.end(), sc.subs.begin(), sc.subs.end());
        pict_off.push_back(pict.size());
        pict.insert(pict.end(), sc.pict.begin(), sc.pict.end());
        pict.resize((pict.size() + 15) & ~(size_t)15);
        uses_off.push_back(uses.size() / 24);
        uses.insert(uses.end(), sc.uses.begin(), sc.uses.end());
        grads_off.push_back(grads.size());
        grads.insert(grads.end(), sc.grads.begin(), sc.grads.end());
        refs_off.push_back(refs.size());
        refs.insert(refs.end(), j.refs.begin(), j.refs.end());
        bounds.insert(bounds.end(), xr.begin(), xr.begin() + 4 * j.refs.size());
        curves_off.push_back(curves.size());
        curves.insert(curves.end(), g.curves.begin(), g.curves.end());
        a->points_off.push_back(points.size() / 2);
        points.insert(points.end(), sc.pairs.begin(), sc.pairs.end());
        const int32_t tile_base = (int32_t)tiles.size();
        std::vector<uint32_t> w;
        build_work(ctx->spec, j.tiles.data(), (int)j.tiles.size(), sc.width, sc.height, w);
        a->work_begin.push_back((int32_t)work.size());
        a->work_count.push_back((int32_t)w.size());
        for (uint32_t u : w) work.push_back(u + ((uint32_t)tile_base << 2));
        tiles.insert(tiles.end(), j.tiles.begin(), j.tiles.end());
        tile_job.insert(tile_job.end(), tj.begin(), tj.end());
        tile_scene.insert(tile_scene.end(), j.tiles.size(), s);
        a->pix_off.push_back(pix);
        pix += (size_t)sc.width * sc.height;
    }
    if (tiles.size() >= (1u << 30)) return set_error(ctx, GUDNI_ERR_ARG, "adjoint: too many tiles");
    a->n_points = points.size() / 2;
    a->n_subs = subs.size();
    a->n_pix = pix;
    a->n_work = (int32_t)work.size();
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    int rc = ensure_scratch(ctx);
    if (rc) return rc;
    rc = ensure_control(ctx, a->n_work);
    if (rc) return rc;
    CHECK_HIP(ctx, upload(ctx, a->geo, geo.data(), geo.size()));
    CHECK_HIP(ctx, upload(ctx, a->subs, subs.data(), 16 * subs.size()));
    CHECK_HIP(ctx, upload(ctx, a->pict, pict.data(), pict.size()));
    CHECK_HIP(ctx, upload(ctx, a->uses, uses.data(), uses.size()));
    CHECK_HIP(ctx, upload(ctx, a->shapes, refs.data(), 16 * refs.size()));
    CHECK_HIP(ctx, upload(ctx, a->bounds, bounds.data(), 4 * bounds.size()));
    CHECK_HIP(ctx, upload(ctx, a->curves, curves.data(), sizeof(CurveSrc) * curves.size()));
    CHECK_HIP(ctx, upload(ctx, a->points, points.data(), 4 * points.size()));
    CHECK_HIP(ctx, upload(ctx, a->tiles, tiles.data(), 32 * tiles.size()));
    CHECK_HIP(ctx, upload(ctx, a->tile_job, tile_job.data(), 4 * tile_job.size()));
    CHECK_HIP(ctx, upload(ctx, a->tile_scene, tile_scene.data(), 4 * tile_scene.size()));
    CHECK_HIP(ctx, upload(ctx, a->work, work.data(), 4 * work.size()));
    CHECK_HIP(ctx, upload(ctx, a->grads, grads.data(), sizeof(gudni_gradient) * grads.size()));
    CHECK_HIP(ctx, a->image.ensure(16 * pix));
    CHECK_HIP(ctx, a->gpix.ensure(16 * pix));
    CHECK_HIP(ctx, a->d_points.ensure(8 * a->n_points));
    CHECK_HIP(ctx, a->d_rgba.ensure(16 * a->n_subs));
    CHECK_HIP(ctx, a->loss.ensure(8 * (size_t)n));
    rc = fill_descs(ctx, a);
    if (rc) return rc;
    // scratch: the raster passes' rings plus a ring of line origins
    if (!a->blocks) {
        int per_cu = 0;
        CHECK_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adj::adjoint_pass, 256, 0));
        a->blocks = std::max(1, std::min(ctx->fast_blocks, ctx->cus * std::max(1, per_cu)));
    }
    const size_t fast_slots = (size_t)a->blocks * 4, slow_slots = (size_t)ctx->slow_blocks * 4;
    size_t scap = 1;
    while ((int)scap < ctx->spec.max_thresholds) scap <<= 1;
    CHECK_HIP(ctx, a->orig_fast.ensure(fast_slots * ((size_t)1 << dev::kFastCapLog2) * 2 * 64 * 16));
    CHECK_HIP(ctx, a->copy_thr.ensure(fast_slots * ((size_t)1 << dev::kFastCapLog2) * 64 * 16));
    CHECK_HIP(ctx, a->copy_hdr.ensure(fast_slots * ((size_t)1 << dev::kFastCapLog2) * 64 * 4));
    CHECK_HIP(ctx, a->ends_fast.ensure(fast_slots * ((size_t)1 << dev::kFastCapLog2) * 4 * 64 * 16));
    CHECK_HIP(ctx, a->orig_slow.ensure(slow_slots * scap * 2 * 64 * 16));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GUDNI_OK;
}

// Scene descriptors: the geometry, tile references, their extents and the
// curve provenance of each scene come from its device rebuild once there is
// one, else from the host build's concatenated buffers.
int fill_descs(gudni_ctx *ctx, gudni_adjoint_batch *a)
{
    const int n = (int)a->scenes.size();
    std::vector<adj::AdjDesc> descs((size_t)n);
    for (int s = 0; s < n; ++s) {
        const gudni_adjoint_batch::Scene &sc = a->scenes[s];
        adj::AdjDesc &d = descs[s];
        std::memset(&d, 0, sizeof(d));
        if (!a->dev.empty()) {
            const gudni_adjoint_batch::DevScene &v = a->dev[s];
            d.s.geo = (const uint8_t *)v.geo.p;
            d.s.shapes = (const ShapeRec *)v.refs.p;
            d.s.bounds = (const float4 *)v.ref_bounds.p;
            d.curves = (const CurveSrc *)v.curves.p;
        } else {
            d.s.geo = (const uint8_t *)a->geo.p + a->geo_off[s];
            d.s.shapes = (const ShapeRec *)a->shapes.p + a->refs_off[s];
            d.s.bounds = (const float4 *)a->bounds.p + a->refs_off[s];
            d.curves = (const CurveSrc *)a->curves.p + a->curves_off[s];
        }
        d.s.substances = (const float4 *)a->subs.p + a->subs_off[s];
        d.s.pict = (const uint8_t *)a->pict.p + a->pict_off[s];
        d.s.uses = (const PictUse *)a->uses.p + a->uses_off[s];
        d.s.out = nullptr;
        d.s.out_f32 = (float4 *)a->image.p + a->pix_off[s];
        d.s.grads = (const float4 *)((const gudni_gradient *)a->grads.p + a->grads_off[s]);
        d.s.filter = ctx->spec.texture_filter;
        d.s.sort_bits = ctx->spec.sort_radix_bits;
        for (int i = 0; i < 4; ++i) d.s.bg[i] = sc.bg[i];
        d.s.width = sc.width;
        d.s.height = sc.height;
        d.points = (const float2 *)a->points.p + a->points_off[s];
        d.target = (const float4 *)a->target.p + a->pix_off[s];
        d.gpix = (float4 *)a->gpix.p + a->pix_off[s];
        d.d_points = (float *)a->d_points.p + 2 * a->points_off[s];
        d.d_rgba = (float *)a->d_rgba.p + 4 * a->subs_off[s];
        d.loss = (double *)a->loss.p + s;
        d.n_points = (int32_t)(sc.pairs.size() / 2);
    }
    CHECK_HIP(ctx, upload(ctx, a->descs, descs.data(), sizeof(adj::AdjDesc) * descs.size()));
    return GUDNI_OK;
}

// Serialises (with curve provenance) and bins scene s on the device from the
// parameters in a->points, into a->dev[s].
// Scene s's serialiser inputs (its outline offsets go to the device once).
int scene_ser_in(gudni_ctx *ctx, gudni_adjoint_batch *a, int s, SerIn &in)
{
    const gudni_adjoint_batch::Scene &sc = a->scenes[s];
    gudni_adjoint_batch::DevScene &v = a->dev[s];
    const int32_t n = (int32_t)sc.tags.size();
    if (!v.tags.p) {
        CHECK_HIP(ctx, upload(ctx, v.tags, sc.tags.data(), 8u * (size_t)n));
        CHECK_HIP(ctx, upload(ctx, v.soff, sc.shape_off.data(), 4u * sc.shape_off.size()));
        CHECK_HIP(ctx, upload(ctx, v.ooff, sc.outline_off.data(), 4u * sc.outline_off.size()));
    }
    in.tags = (const uint64_t *)v.tags.p;
    in.soff = (const int32_t *)v.soff.p;
    in.ooff = (const int32_t *)v.ooff.p;
    in.pairs = (const float *)a->points.p + 2 * a->points_off[s];
    in.n_shapes = n;
    in.width = sc.width;
    in.height = sc.height;
    in.subs = (const float4 *)a->subs.p + a->subs_off[s];
    in.n_subs = (int32_t)sc.subs.size();
    in.n_uses = (int32_t)(sc.uses.size() / 24);
    in.n_grads = (int32_t)sc.grads.size();
    return GUDNI_OK;
}

// Re-serialises the listed scenes from the parameters on the device (one
// scene alone, several in one many-scene pass), with curve provenance.
int device_serialize_scenes(gudni_ctx *ctx, gudni_adjoint_batch *a, const std::vector<int> &which)
{
    const size_t m = which.size();
    std::vector<SerIn> in(m);
    std::vector<SerOut> out(m);
    std::vector<SerCounts> cnt(m);
    for (size_t i = 0; i < m; ++i) {
        const int rc = scene_ser_in(ctx, a, which[i], in[i]);
        if (rc) return rc;
        gudni_adjoint_batch::DevScene &v = a->dev[which[i]];
        out[i] = SerOut{&v.geo, &v.shapes, &v.bbox, &v.sbounds, &v.curves};
    }
    const int rc = m == 1 ? serialize_outlines(ctx, in[0], out[0], cnt[0])
                          : serialize_outlines_many(ctx, in.data(), out.data(), cnt.data(), (int)m);
    if (rc) return rc;
    for (size_t i = 0; i < m; ++i) a->dev[which[i]].n_kept = (int32_t)cnt[i].n_kept;
    return GUDNI_OK;
}

// Bins the listed scenes' serialised shapes: one scene with the device-wide
// level launches (a whole GPU on one tree), several in one launch of a block
// per scene (the level launches would cost more than the work).
int device_bin_scenes(gudni_ctx *ctx, gudni_adjoint_batch *a, const std::vector<int> &which)
{
    const size_t n = which.size();
    if (n == 1) {
        const int s = which[0];
        gudni_adjoint_batch::DevScene &v = a->dev[s];
        const BinIn in{(const ShapeRec *)v.shapes.p, (const float4 *)v.bbox.p, (const float4 *)v.sbounds.p,
                       v.n_kept, a->scenes[s].width, a->scenes[s].height};
        const BinOut out{&v.tiles, &v.tile_job, &v.refs, &v.ref_bounds, &v.work};
        BinCounts c;
        const int rc = bin_scene(ctx, in, out, c);
        if (rc) return rc;
        v.n_tiles = c.n_tiles;
        v.n_work = c.n_work;
        return GUDNI_OK;
    }
    std::vector<BinIn> in(n);
    std::vector<BinOut> out(n);
    std::vector<BinCounts> cnt(n);
    for (size_t i = 0; i < n; ++i) {
        const int s = which[i];
        gudni_adjoint_batch::DevScene &v = a->dev[s];
        in[i] = BinIn{(const ShapeRec *)v.shapes.p, (const float4 *)v.bbox.p, (const float4 *)v.sbounds.p,
                      v.n_kept, a->scenes[s].width, a->scenes[s].height};
        out[i] = BinOut{&v.tiles, &v.tile_job, &v.refs, &v.ref_bounds, &v.work};
    }
    const int rc = bin_scenes(ctx, in.data(), out.data(), cnt.data(), (int)n);
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) {
        a->dev[which[i]].n_tiles = cnt[i].n_tiles;
        a->dev[which[i]].n_work = cnt[i].n_work;
    }
    return GUDNI_OK;
}

// The batch's tile, job, scene and work tables from every scene's device
// rebuild, and the descriptors pointing at them.
int assemble(gudni_ctx *ctx, gudni_adjoint_batch *a)
{
    const hipStream_t st = ctx->stream;
    const int n = (int)a->scenes.size();
    size_t T = 0, W = 0;
    for (const auto &v : a->dev) { T += (size_t)v.n_tiles; W += (size_t)v.n_work; }
    if (T >= (1u << 30)) return set_error(ctx, GUDNI_ERR_ARG, "adjoint: too many tiles");
    CHECK_HIP(ctx, a->tiles.ensure(32 * T));
    CHECK_HIP(ctx, a->tile_job.ensure(4 * T));
    CHECK_HIP(ctx, a->tile_scene.ensure(4 * T));
    CHECK_HIP(ctx, a->work.ensure(4 * W));
    a->work_begin.assign((size_t)n, 0);
    a->work_count.assign((size_t)n, 0);
    size_t tb = 0, wb = 0;
    for (int s = 0; s < n; ++s) {
        const gudni_adjoint_batch::DevScene &v = a->dev[s];
        if (v.n_tiles) {
            CHECK_HIP(ctx, hipMemcpyAsync((uint8_t *)a->tiles.p + 32 * tb, v.tiles.p, 32 * (size_t)v.n_tiles,
                                          hipMemcpyDeviceToDevice, st));
            CHECK_HIP(ctx, hipMemcpyAsync((int32_t *)a->tile_job.p + tb, v.tile_job.p, 4 * (size_t)v.n_tiles,
                                          hipMemcpyDeviceToDevice, st));
            hipLaunchKernelGGL(adj::adj_fill_i32, dim3(std::min(1024, (v.n_tiles + 255) / 256)), dim3(256), 0, st,
                               (int32_t *)a->tile_scene.p + tb, v.n_tiles, (int32_t)s);
        }
        if (v.n_work)
            hipLaunchKernelGGL(adj::adj_rebase_work, dim3(std::min(1024, (v.n_work + 255) / 256)), dim3(256), 0, st,
                               (const uint32_t *)v.work.p, (uint32_t *)a->work.p + wb, v.n_work, (uint32_t)tb);
        a->work_begin[s] = (int32_t)wb;
        a->work_count[s] = v.n_work;
        tb += (size_t)v.n_tiles;
        wb += (size_t)v.n_work;
    }
    CHECK_HIP(ctx, hipGetLastError());
    a->n_work = (int32_t)wb;
    int rc = ensure_control(ctx, a->n_work);
    if (rc) return rc;
    return fill_descs(ctx, a);
}

// After new outline points for `scene`: the first time every scene is
// rebuilt on the device (so all tables come from one place), afterwards
// only that scene.
int device_update(gudni_ctx *ctx, gudni_adjoint_batch *a, int scene)
{
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    // a launch may still read the per-scene buffers the rebuild can regrow
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int rc = GUDNI_OK;
    std::vector<int> which;
    if (a->dev.empty() || scene < 0) {
        a->dev.resize(a->scenes.size());
        for (int s = 0; s < (int)a->scenes.size(); ++s) which.push_back(s);
    } else {
        which.push_back(scene);
    }
    // any failure from here on leaves the parameters on the device without
    // a complete set of tables (a regrown buffer may be gone, counts stale):
    // the batch refuses to run until it is uploaded again (or every scene
    // is rebuilt by a successful update)
    const bool was_broken = a->broken;
    a->broken = true;
    rc = device_serialize_scenes(ctx, a, which);
    if (rc == GUDNI_OK) rc = device_bin_scenes(ctx, a, which);
    if (rc == GUDNI_OK) rc = assemble(ctx, a);
    if (rc) return rc;
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    a->broken = was_broken && which.size() < a->scenes.size();
    return GUDNI_OK;
}

int copy_scene(gudni_ctx *ctx, const gudni_adjoint_scene &in, gudni_adjoint_batch::Scene &sc, int s)
{
    const gudni_outlines &o = in.outlines;
    if (in.width <= 0 || in.height <= 0 || o.n_shapes < 0 || in.n_substances < 0 || in.n_pict_uses < 0 ||
        (o.n_shapes && (!o.tags || !o.shape_outline_offsets || !o.outline_pair_offsets)) ||
        (in.n_substances && !in.rgba) || (in.pict_bytes && !in.pict_data) ||
        (in.n_pict_uses && !in.pict_uses))
        return set_error(ctx, GUDNI_ERR_ARG, "adjoint scene " + std::to_string(s) + ": bad argument");
    int rc = validate_uses(ctx, in.pict_uses, in.n_pict_uses, in.pict_bytes);
    if (rc) return rc;
    rc = validate_gradients(ctx, in.gradients, in.n_gradients);
    if (rc) return rc;
    {
        int32_t n_outlines = 0, n_pairs = 0;   // the device serialiser indexes with these
        rc = check_outline_offsets(ctx, &o, n_outlines, n_pairs);
        if (rc) return set_error(ctx, rc, "adjoint scene " + std::to_string(s) + ": " + ctx->error);
    }
    sc.width = in.width;
    sc.height = in.height;
    std::memcpy(sc.bg, in.background, sizeof(sc.bg));
    sc.tags.assign(o.tags, o.tags + o.n_shapes);
    sc.shape_off.assign(o.shape_outline_offsets, o.shape_outline_offsets + o.n_shapes + 1);
    const int n_outlines = o.n_shapes ? sc.shape_off.back() : 0;
    if (n_outlines < 0) return set_error(ctx, GUDNI_ERR_ARG, "adjoint: bad outline offsets");
    sc.outline_off.assign(o.outline_pair_offsets, o.outline_pair_offsets + n_outlines + 1);
    const int n_pairs = n_outlines ? sc.outline_off.back() : 0;
    if (n_pairs < 0 || (n_pairs && !o.pairs)) return set_error(ctx, GUDNI_ERR_ARG, "adjoint: bad pairs");
    sc.pairs.assign(o.pairs, o.pairs + 4 * (size_t)n_pairs);
    sc.subs.resize((size_t)in.n_substances);
    if (in.n_substances) std::memcpy(sc.subs.data(), in.rgba, 16 * (size_t)in.n_substances);
    sc.pict.assign((const uint8_t *)in.pict_data, (const uint8_t *)in.pict_data + in.pict_bytes);
    sc.uses.assign((const uint8_t *)in.pict_uses, (const uint8_t *)in.pict_uses + 24 * (size_t)in.n_pict_uses);
    sc.grads.assign((const gudni_gradient *)in.gradients,
                    (const gudni_gradient *)in.gradients + (in.gradients ? in.n_gradients : 0));
    return check_scene(ctx, sc, s);
}

int launch(gudni_ctx *ctx, gudni_adjoint_batch *a, int32_t work_begin, int32_t n_work, size_t clear_pts0,
           size_t clear_pts1, size_t clear_subs0, size_t clear_subs1, int scene0, int scene1)
{
    const hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(adj::adj_clear, dim3(256), dim3(256), 0, st, (float *)a->d_points.p + 2 * clear_pts0,
                       2 * (clear_pts1 - clear_pts0));
    hipLaunchKernelGGL(adj::adj_clear, dim3(64), dim3(256), 0, st, (float *)a->d_rgba.p + 4 * clear_subs0,
                       4 * (clear_subs1 - clear_subs0));
    hipLaunchKernelGGL(adj::adj_clear, dim3(1), dim3(64), 0, st, (float *)a->loss.p + 2 * scene0,
                       (size_t)2 * (scene1 - scene0));
    uint32_t *ctl = (uint32_t *)ctx->control.p;
    hipLaunchKernelGGL(adj::adj_reset, dim3(1), dim3(64), 0, st, ctl, dev::kCtlWords);
    CHECK_HIP(ctx, hipGetLastError());
    adj::AdjFrame af;
    af.f = base_frame(ctx);
    af.f.coverage_mode = GUDNI_COVERAGE_ACCURATE;
    af.f.stage_cycles = nullptr;
    if (ctx->timing) {
        CHECK_HIP(ctx, ctx->cycles.ensure(8 * sizeof(unsigned long long)));
        hipLaunchKernelGGL(adj::adj_reset, dim3(1), dim3(64), 0, st, (uint32_t *)ctx->cycles.p, 16);
        af.f.stage_cycles = (unsigned long long *)ctx->cycles.p;
        ctx->timed_adjoint = true;
    }
    af.f.stochastic = 0.0f;
    af.f.scenes = nullptr;
    af.f.tiles = (const TileRec *)a->tiles.p;
    af.f.tile_job = (const int32_t *)a->tile_job.p;
    af.f.tile_scene = (const int32_t *)a->tile_scene.p;
    af.f.work = (const uint32_t *)a->work.p + work_begin;
    af.f.n_work = n_work;
    af.f.status = ctl + dev::kCtlStatus;
    ctx->status_word = af.f.status;
    af.f.retry = ctl + dev::kCtlRetry;
    af.f.retry_capacity = ctx->retry_capacity;
    af.scenes = (const adj::AdjDesc *)a->descs.p;
    af.f.counter = ctl + dev::kCtlFast;
    af.f.n_parts = dev::kParts;
    af.f.q_thr = (float4 *)ctx->fast_thr.p;
    af.f.q_hdr = (uint32_t *)ctx->fast_hdr.p;
    af.f.shape_idx = (int32_t *)ctx->fast_sidx.p;
    af.f.cap_log2 = dev::kFastCapLog2;
    af.q_orig = (float4 *)a->orig_fast.p;
    af.copy_thr = (float4 *)a->copy_thr.p;
    af.copy_hdr = (uint32_t *)a->copy_hdr.p;
    af.q_ends = (float4 *)a->ends_fast.p;
    const int grid = std::max(1, std::min(a->blocks, (n_work + 3) / 4));
    hipLaunchKernelGGL(adj::adjoint_pass, dim3(grid), dim3(256), 0, st, af, 0);
    CHECK_HIP(ctx, hipGetLastError());
    int slog = 0;
    while ((1 << slog) < ctx->spec.max_thresholds) ++slog;
    af.f.counter = ctl + dev::kCtlSlow;
    af.f.n_parts = 1;
    af.f.q_thr = (float4 *)ctx->slow_thr.p;
    af.f.q_hdr = (uint32_t *)ctx->slow_hdr.p;
    af.f.shape_idx = (int32_t *)ctx->slow_sidx.p;
    af.f.cap_log2 = slog;
    af.q_orig = (float4 *)a->orig_slow.p;
    af.copy_thr = nullptr;   // the retry pass's rings are large: it regenerates instead
    af.copy_hdr = nullptr;
    af.q_ends = nullptr;
    hipLaunchKernelGGL(adj::adjoint_pass, dim3(ctx->slow_blocks), dim3(256), 0, st, af, 1);
    CHECK_HIP(ctx, hipGetLastError());
    return GUDNI_OK;
}

}  // namespace

extern "C" {

int gudni_adjoint_upload(gudni_ctx *ctx, int32_t n_scenes, const gudni_adjoint_scene *scenes,
                         gudni_adjoint_batch **out)
{
    if (!ctx || !scenes || !out || n_scenes <= 0) return GUDNI_ERR_ARG;
    *out = nullptr;
    gudni_adjoint_batch *a = new gudni_adjoint_batch();
    a->device = ctx->device;
    a->scenes.resize((size_t)n_scenes);
    int rc = GUDNI_OK;
    std::vector<float> target;
    for (int s = 0; s < n_scenes && rc == GUDNI_OK; ++s) {
        rc = copy_scene(ctx, scenes[s], a->scenes[s], s);
        if (rc == GUDNI_OK && !scenes[s].target)
            rc = set_error(ctx, GUDNI_ERR_ARG, "adjoint scene " + std::to_string(s) + ": no target");
    }
    if (rc == GUDNI_OK) {
        size_t pix = 0;
        for (int s = 0; s < n_scenes; ++s) pix += (size_t)scenes[s].width * scenes[s].height;
        target.resize(4 * pix);
        size_t off = 0;
        for (int s = 0; s < n_scenes; ++s) {
            const size_t np = (size_t)scenes[s].width * scenes[s].height;
            std::memcpy(target.data() + 4 * off, scenes[s].target, 16 * np);
            off += np;
        }
        const hipError_t e = upload(ctx, a->target, target.data(), 4 * target.size());
        if (e != hipSuccess) rc = set_error(ctx, GUDNI_ERR_HIP, std::string("adjoint: ") + hipGetErrorString(e));
    }
    if (rc == GUDNI_OK) rc = build(ctx, a);
    if (rc) {
        gudni_adjoint_free(a);
        return rc;
    }
    *out = a;
    return GUDNI_OK;
}

int gudni_adjoint_update(gudni_ctx *ctx, gudni_adjoint_batch *a, int32_t scene, const float *pairs,
                         const float *rgba)
{
    if (!ctx || !a || scene < -1 || scene >= (int32_t)a->scenes.size()) return GUDNI_ERR_ARG;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int s0 = scene < 0 ? 0 : scene, s1 = scene < 0 ? (int)a->scenes.size() : scene + 1;
    for (int s = s0; s < s1; ++s) {   // with scene -1 the inputs are every scene's, concatenated
        gudni_adjoint_batch::Scene &sc = a->scenes[s];
        if (pairs) {
            const float *src = pairs + (scene < 0 ? 2 * a->points_off[s] : 0);
            std::memcpy(sc.pairs.data(), src, 4 * sc.pairs.size());
            CHECK_HIP(ctx, hipMemcpyAsync((float *)a->points.p + 2 * a->points_off[s], sc.pairs.data(),
                                          4 * sc.pairs.size(), hipMemcpyHostToDevice, ctx->stream));
        }
        if (rgba) {
            const float *src = rgba + (scene < 0 ? 4 * a->subs_off[s] : 0);
            std::memcpy(sc.subs.data(), src, 16 * sc.subs.size());
            CHECK_HIP(ctx, hipMemcpyAsync((float4 *)a->subs.p + a->subs_off[s], sc.subs.data(),
                                          16 * sc.subs.size(), hipMemcpyHostToDevice, ctx->stream));
        }
    }
    if (pairs) return device_update(ctx, a, scene);
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GUDNI_OK;
}

int gudni_adjoint_update_device(gudni_ctx *ctx, gudni_adjoint_batch *a, int32_t scene, const float *pairs,
                                const float *rgba)
{
    if (!ctx || !a || scene < -1 || scene >= (int32_t)a->scenes.size()) return GUDNI_ERR_ARG;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    if (scene < 0) {   // every scene's parameters, concatenated in batch order
        if (pairs && a->n_points)
            CHECK_HIP(ctx, hipMemcpyAsync(a->points.p, pairs, 8 * a->n_points, hipMemcpyDeviceToDevice,
                                          ctx->stream));
        if (rgba && a->n_subs)
            CHECK_HIP(ctx, hipMemcpyAsync(a->subs.p, rgba, 16 * a->n_subs, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        const gudni_adjoint_batch::Scene &sc = a->scenes[scene];
        if (pairs && !sc.pairs.empty())
            CHECK_HIP(ctx, hipMemcpyAsync((float *)a->points.p + 2 * a->points_off[scene], pairs,
                                          4 * sc.pairs.size(), hipMemcpyDeviceToDevice, ctx->stream));
        if (rgba && !sc.subs.empty())
            CHECK_HIP(ctx, hipMemcpyAsync((float4 *)a->subs.p + a->subs_off[scene], rgba, 16 * sc.subs.size(),
                                          hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (pairs) return device_update(ctx, a, scene);
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GUDNI_OK;
}

int gudni_adjoint_run(gudni_ctx *ctx, gudni_adjoint_batch *a)
{
    if (!ctx || !a) return GUDNI_ERR_ARG;
    if (a->broken) return set_error(ctx, GUDNI_ERR_ARG, "adjoint batch: an update failed; upload it again");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    return launch(ctx, a, 0, a->n_work, 0, a->n_points, 0, a->n_subs, 0, (int)a->scenes.size());
}

int gudni_adjoint_run_scene(gudni_ctx *ctx, gudni_adjoint_batch *a, int32_t scene)
{
    if (!ctx || !a || scene < 0 || scene >= (int32_t)a->scenes.size()) return GUDNI_ERR_ARG;
    if (a->broken) return set_error(ctx, GUDNI_ERR_ARG, "adjoint batch: an update failed; upload it again");
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    const size_t p0 = a->points_off[scene], p1 = p0 + a->scenes[scene].pairs.size() / 2;
    const size_t s0 = a->subs_off[scene], s1 = s0 + a->scenes[scene].subs.size();
    return launch(ctx, a, a->work_begin[scene], a->work_count[scene], p0, p1, s0, s1, scene, scene + 1);
}

int gudni_adjoint_result(gudni_ctx *ctx, gudni_adjoint_batch *a, int32_t scene, double *loss, float *d_pairs,
                         float *d_rgba, float *image)
{
    if (!ctx || !a || scene < 0 || scene >= (int32_t)a->scenes.size()) return GUDNI_ERR_ARG;
    int rc = read_status(ctx);
    if (rc) return rc;
    if (ctx->timed_adjoint) {
        CHECK_HIP(ctx, hipMemcpy(ctx->last_cycles, ctx->cycles.p, sizeof(ctx->last_cycles),
                                 hipMemcpyDeviceToHost));
        ctx->timed_adjoint = false;
    }
    const gudni_adjoint_batch::Scene &sc = a->scenes[scene];
    if (loss)
        CHECK_HIP(ctx, hipMemcpy(loss, (double *)a->loss.p + scene, 8, hipMemcpyDeviceToHost));
    if (d_pairs && !sc.pairs.empty())
        CHECK_HIP(ctx, hipMemcpy(d_pairs, (float *)a->d_points.p + 2 * a->points_off[scene],
                                 4 * sc.pairs.size(), hipMemcpyDeviceToHost));
    if (d_rgba && !sc.subs.empty())
        CHECK_HIP(ctx, hipMemcpy(d_rgba, (float *)a->d_rgba.p + 4 * a->subs_off[scene],
                                 16 * sc.subs.size(), hipMemcpyDeviceToHost));
    if (image)
        CHECK_HIP(ctx, hipMemcpy(image, (float4 *)a->image.p + a->pix_off[scene],
                                 16 * (size_t)sc.width * sc.height, hipMemcpyDeviceToHost));
    return GUDNI_OK;
}

int gudni_adjoint_result_device(gudni_ctx *ctx, gudni_adjoint_batch *a, int32_t scene, double *loss,
                                float *d_pairs, float *d_rgba, float *image)
{
    if (!ctx || !a || scene < -1 || scene >= (int32_t)a->scenes.size()) return GUDNI_ERR_ARG;
    if (scene < 0) {   // every scene: losses [n_scenes], gradients and images concatenated
        const hipStream_t st = ctx->stream;
        CHECK_HIP(ctx, hipSetDevice(ctx->device));
        if (loss)
            CHECK_HIP(ctx, hipMemcpyAsync(loss, a->loss.p, 8 * a->scenes.size(), hipMemcpyDeviceToDevice, st));
        if (d_pairs && a->n_points)
            CHECK_HIP(ctx, hipMemcpyAsync(d_pairs, a->d_points.p, 8 * a->n_points, hipMemcpyDeviceToDevice, st));
        if (d_rgba && a->n_subs)
            CHECK_HIP(ctx, hipMemcpyAsync(d_rgba, a->d_rgba.p, 16 * a->n_subs, hipMemcpyDeviceToDevice, st));
        if (image && a->n_pix)
            CHECK_HIP(ctx, hipMemcpyAsync(image, a->image.p, 16 * a->n_pix, hipMemcpyDeviceToDevice, st));
        return GUDNI_OK;
    }
    const gudni_adjoint_batch::Scene &sc = a->scenes[scene];
    const hipStream_t st = ctx->stream;
    CHECK_HIP(ctx, hipSetDevice(ctx->device));
    if (loss)
        CHECK_HIP(ctx, hipMemcpyAsync(loss, (double *)a->loss.p + scene, 8, hipMemcpyDeviceToDevice, st));
    if (d_pairs && !sc.pairs.empty())
        CHECK_HIP(ctx, hipMemcpyAsync(d_pairs, (float *)a->d_points.p + 2 * a->points_off[scene],
                                      4 * sc.pairs.size(), hipMemcpyDeviceToDevice, st));
    if (d_rgba && !sc.subs.empty())
        CHECK_HIP(ctx, hipMemcpyAsync(d_rgba, (float *)a->d_rgba.p + 4 * a->subs_off[scene],
                                      16 * sc.subs.size(), hipMemcpyDeviceToDevice, st));
    if (image)
        CHECK_HIP(ctx, hipMemcpyAsync(image, (float4 *)a->image.p + a->pix_off[scene],
                                      16 * (size_t)sc.width * sc.height, hipMemcpyDeviceToDevice, st));
    return GUDNI_OK;
}

int gudni_adjoint_check(gudni_ctx *ctx, gudni_adjoint_batch *a)
{
    if (!ctx || !a) return GUDNI_ERR_ARG;
    return read_status(ctx);
}

void gudni_adjoint_free(gudni_adjoint_batch *a)
{
    if (!a) return;
    (void)hipSetDevice(a->device);
    DevBuf *bufs[] = {&a->geo, &a->subs, &a->pict, &a->uses, &a->shapes, &a->bounds, &a->curves,
                      &a->points, &a->target, &a->image, &a->gpix, &a->d_points, &a->d_rgba, &a->loss,
                      &a->tiles, &a->tile_job, &a->tile_scene, &a->work, &a->descs, &a->orig_fast,
                      &a->orig_slow, &a->grads, &a->copy_thr, &a->copy_hdr, &a->ends_fast};
    for (DevBuf *b : bufs) b->release();
    a->drop_device_scenes();
    delete a;
}

int gudni_adjoint(gudni_ctx *ctx, const gudni_adjoint_scene *scene, double *loss, float *d_pairs,
                  float *d_rgba, float *image)
{
    gudni_adjoint_batch *a = nullptr;
    int rc = gudni_adjoint_upload(ctx, 1, scene, &a);
    if (rc) return rc;
    rc = gudni_adjoint_run(ctx, a);
    if (rc == GUDNI_OK) rc = gudni_adjoint_result(ctx, a, 0, loss, d_pairs, d_rgba, image);
    gudni_adjoint_free(a);
    return rc;
}

}  // extern "C"
