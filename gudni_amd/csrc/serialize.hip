// serialize.hip -- the onShape serialiser on the device: closed outlines of
// curve pairs -> knob-free quadratics -> x-monotone strands of at most 16
// curves in Eytzinger order -> the strand heap, plus one Shape record, box
// and strand extent per kept shape.  The steps are those of the host
// serialiser (host.cpp; Raster/Strand.hs:77-206, Deknob.hs:61-108,
// ReorderTable.hs:44-111, Serialize.hs:76-177, Outline.hs:123-133) with the
// same FP32 operations in the same order, so the bytes are the same.
//
// Layout of the work: one thread per shape (shapes hold a few outlines of
// 3..30 pairs; the walk over a shape's curves is short and serial).
//   count     box, cull, substance checks, and per outline a streaming pass
//             over its knob-split curves that sizes every run of one x
//             direction (16-byte units and strands) in O(1) state
//   scans     kept-shape numbering and heap offsets (exclusive scans)
//   emit      the same pass again, buffering at most 16 curves, writes each
//             strand where it belongs (the run still open when an outline's
//             loop ends is written first, Strand.hs:77-102), then the
//             shape's record and box
//   extents   per kept shape the strands' extent (the generation prefilter's
//             input), read back from the heap as the host computes it
// With curve tracking (the adjoint's serialisation) each strand's header
// dword 1 holds the index of its first curve and every strand curve gets a
// CurveSrc record (its outline points and knob-split half), numbered in heap
// order, as the host's tracked build numbers them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"
#include "device.h"
#include "scan.h"
#include "serialize.h"

namespace gudni {
namespace ser {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kErrSubstance = 1u, kErrGradient = 2u, kErrPicture = 4u;

struct Pt { float x, y; };
struct Quad { Pt p0, c, p1; };
struct Orders { int8_t o[kMaxStrandCurves + 1][kMaxStrandCurves]; };

struct Args {
    const uint64_t *tags;
    const int32_t *soff, *ooff;
    const float *pairs;
    int32_t n_shapes;
    float fw, fh;
    const float4 *subs;
    int32_t n_subs, n_uses, n_grads;
    uint32_t *keep, *strands, *status;
    uint64_t *units, *kidx, *ustart;
    float4 *box;
    uint8_t *heap;
    ShapeRec *shapes;
    float4 *bbox, *bounds;
    CurveSrc *curves;           // null: no curve tracking
    uint64_t *cstart;           // per shape: its first curve (scan)
};

// between t v0 v1 = v0 ^* (1-t) ^+^ v1 ^* t   (Deknob.hs:40-41)
__device__ __forceinline__ Pt lerp_t(float t, Pt a, Pt b)
{
    const float s = 1.0f - t;
    return Pt{a.x * s + b.x * t, a.y * s + b.y * t};
}

// A control point strictly outside the end points' x range is a knob: split
// where dx/dt changes sign, found by bisection on FP32 parameters to
// iota = 1e-4 (Deknob.hs:61-79, Figure/Space.hs:101-102).
__device__ int split_knob(const Quad &q, Quad *out, float &split_t)
{
    const bool bulge_left = q.c.x < q.p0.x && q.c.x < q.p1.x;
    const bool bulge_right = q.c.x > q.p0.x && q.c.x > q.p1.x;
    split_t = 0.0f;
    if (!bulge_left && !bulge_right) {
        out[0] = q;
        return 1;
    }
    float lo = 0.0f, hi = 1.0f, t = 0.5f;
    Pt m0, on, m1;
    for (;;) {
        m0 = lerp_t(t, q.p0, q.c);
        m1 = lerp_t(t, q.c, q.p1);
        on = lerp_t(t, m0, m1);
        if (hi - lo <= 0.0001f) break;
        const bool out1 = bulge_left ? m1.x < on.x : m1.x > on.x;
        const bool out0 = bulge_left ? m0.x < on.x : m0.x > on.x;
        if (out1) {                // still bulging: move toward p1
            const float nt = t + ((hi - t) / 2.0f);
            lo = t;
            t = nt;
        } else if (out0) {         // overshot: move toward p0
            const float nt = lo + ((t - lo) / 2.0f);
            hi = t;
            t = nt;
        } else {
            break;
        }
    }
    out[0] = Quad{q.p0, m0, on};
    out[1] = Quad{on, m1, q.p1};
    split_t = t;
    return 2;
}

__device__ __forceinline__ int x_dir(const Quad &q)
{
    return (q.p0.x < q.p1.x) ? -1 : (q.p0.x > q.p1.x ? 1 : 0);
}

// Curve i of a closed outline of k pairs: (on_i, off_i, on_{i+1 mod k}).
__device__ __forceinline__ Quad curve_at(const float *p, int i, int k)
{
    const int j = (i + 1 == k) ? 0 : i + 1;
    return Quad{{p[4 * i], p[4 * i + 1]}, {p[4 * i + 2], p[4 * i + 3]}, {p[4 * j], p[4 * j + 1]}};
}

// 16-byte units of a run of L curves cut into strands of <= 16: each strand
// of m curves takes m + 1.
__device__ __forceinline__ uint32_t run_units(uint32_t L)
{
    return L + (L + kMaxStrandCurves - 1) / kMaxStrandCurves;
}

struct Runs {
    uint32_t units, strands;      // of the whole outline
    uint32_t last_start;          // curve index where the last run starts
    uint32_t last_units;          // its units (written first)
    uint32_t last_len;            // and its curves
};

// Maximal runs of knob-free curves sharing a non-zero x direction (a
// vertical curve is a run of its own).
__device__ Runs count_runs(const float *p, int k)
{
    Runs r{0u, 0u, 0u, 0u, 0u};
    if (k < 2) return r;
    int dir = 0;
    uint32_t len = 0, start = 0, idx = 0;
    for (int i = 0; i < k; ++i) {
        Quad two[2];
        float t;
        const int n = split_knob(curve_at(p, i, k), two, t);
        for (int h = 0; h < n; ++h, ++idx) {
            const int d = x_dir(two[h]);
            if (idx == 0) {
                dir = d;
                len = 1;
            } else if (d != dir || dir == 0) {
                r.units += run_units(len);
                r.strands += (len + kMaxStrandCurves - 1) / kMaxStrandCurves;
                dir = d;
                len = 1;
                start = idx;
            } else {
                ++len;
            }
        }
    }
    r.units += run_units(len);
    r.strands += (len + kMaxStrandCurves - 1) / kMaxStrandCurves;
    r.last_start = start;
    r.last_units = run_units(len);
    r.last_len = len;
    return r;
}

// One strand of m curves (Strand.hs:180-206): header {u16 size in 8-byte
// units, u16 flags (GUDNI_STRAND_REVERSED, nonzero-fill shapes only), u32 0
// or the first curve's index} and the right end point, the left end point and its
// control, then the interior (point, control) pairs in tree order.  A chunk
// that runs right to left is reversed first (reverseIfBackwards).
__device__ void write_strand(const Quad *buf, int m, uint8_t *dst, const Orders &ord, uint32_t curve_base,
                             bool mark_reversed)
{
    const bool rev = buf[0].p0.x > buf[m - 1].p1.x;
    auto at = [&](int q) -> Quad {
        if (!rev) return buf[q];
        const Quad &o = buf[m - 1 - q];
        return Quad{o.p1, o.c, o.p0};
    };
    const Quad first = at(0), last = at(m - 1);
    uint4 h;
    h.x = (uint32_t)(2 * (m + 1)) | ((mark_reversed && rev) ? (uint32_t)GUDNI_STRAND_REVERSED << 16 : 0u);
    h.y = curve_base;   // 0 without tracking
    h.z = __float_as_uint(last.p1.x);
    h.w = __float_as_uint(last.p1.y);
    reinterpret_cast<uint4 *>(dst)[0] = h;
    reinterpret_cast<float4 *>(dst)[1] = make_float4(first.p0.x, first.p0.y, first.c.x, first.c.y);
    for (int k = 0; k < m - 1; ++k) {
        const Quad node = at(1 + ord.o[m][k]);
        reinterpret_cast<float4 *>(dst)[2 + k] = make_float4(node.p0.x, node.p0.y, node.c.x, node.c.y);
    }
}

// Writes one outline's strands from heap unit `base` on; with tracking,
// their curves' records from curve `cbase` on (pair_base: the outline's
// first pair in the scene).
__device__ void emit_outline(const float *p, int k, const Runs &r, uint8_t *heap, uint64_t base,
                             const Orders &ord, CurveSrc *curves, uint64_t cbase, int32_t pair_base,
                             bool mark_reversed)
{
    if (k < 2) return;
    uint64_t pos_last = base, pos_rest = base + r.last_units;
    uint64_t cpos_last = cbase, cpos_rest = cbase + r.last_len;
    Quad buf[kMaxStrandCurves];
    CurveSrc src[kMaxStrandCurves];
    int nb = 0, dir = 0;
    bool in_last = false;
    uint32_t idx = 0;
    auto flush = [&]() {
        if (!nb) return;
        uint64_t &pos = in_last ? pos_last : pos_rest;
        uint64_t &cpos = in_last ? cpos_last : cpos_rest;
        write_strand(buf, nb, heap + 16 * pos, ord, curves ? (uint32_t)cpos : 0u, mark_reversed);
        if (curves) {
            const bool rev = buf[0].p0.x > buf[nb - 1].p1.x;
            for (int q = 0; q < nb; ++q) {
                CurveSrc c = src[rev ? nb - 1 - q : q];
                if (rev) c.flags |= kCurveReversed;
                curves[cpos + q] = c;
            }
        }
        pos += (uint64_t)nb + 1;
        cpos += (uint64_t)nb;
        nb = 0;
    };
    for (int i = 0; i < k; ++i) {
        Quad two[2];
        float t;
        const int n = split_knob(curve_at(p, i, k), two, t);
        const int j = (i + 1 == k) ? 0 : i + 1;
        const int32_t a = 2 * (pair_base + i), b = 2 * (pair_base + j);
        for (int h = 0; h < n; ++h, ++idx) {
            const int d = x_dir(two[h]);
            if (idx == 0) {
                dir = d;
                in_last = r.last_start == 0;
            } else if (d != dir || dir == 0) {
                flush();
                dir = d;
                in_last = idx == r.last_start;
            } else if (nb == kMaxStrandCurves) {
                flush();
            }
            buf[nb] = two[h];
            if (curves)
                src[nb] = CurveSrc{a, a + 1, b, t, n == 2 ? (h == 0 ? kCurveLowerHalf : kCurveUpperHalf) : 0u};
            ++nb;
        }
    }
    flush();
}

__device__ __forceinline__ int gtid() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }
__device__ __forceinline__ int gstride() { return (int)(gridDim.x * blockDim.x); }

// Kernel bodies over (first thread, stride): the single-scene kernels run
// them over the grid, the many-scene kernels over the blocks of a grid row
// (blockIdx.y = scene).
__device__ void count_body(const Args &a, int t0, int ts)
{
    for (int s = t0; s < a.n_shapes; s += ts) {
        const int o0 = a.soff[s], o1 = a.soff[s + 1];
        const int p0 = a.ooff[o0], p1 = a.ooff[o1];
        bool keep = false;
        float l = 0.f, t = 0.f, r = 0.f, b = 0.f;
        if (p1 > p0) {
            // bounding box over every on- and off-curve point (Outline.hs:123-133)
            l = a.pairs[4 * (size_t)p0];
            t = a.pairs[4 * (size_t)p0 + 1];
            r = l;
            b = t;
            for (int p = p0; p < p1; ++p) {
                const float4 q = reinterpret_cast<const float4 *>(a.pairs)[p];
                const float xs[2] = {q.x, q.z}, ys[2] = {q.y, q.w};
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const float x = xs[h], y = ys[h];
                    if (!(l <= x)) l = x;
                    if (!(x <= r)) r = x;
                    if (!(t <= y)) t = y;
                    if (!(y <= b)) b = y;
                }
            }
            keep = !(l >= a.fw || t >= a.fh || r <= 0.0f || b <= 0.0f);   // excludeBox
        }
        uint64_t units = 0;
        uint32_t strands = 0;
        if (keep) {
            const uint64_t tag = a.tags[s];
            const uint64_t sid = tag & dev::kTagIdMask;
            if (sid >= (uint64_t)a.n_subs) {
                atomicOr(a.status, kErrSubstance);
            } else {
                const uint64_t type = tag & dev::kTagTypeMask;
                if (type != dev::kTagSolid) {
                    const uint32_t id = __float_as_uint(a.subs[sid].x);
                    if (type == dev::kTagGradient) {
                        if (id >= (uint32_t)a.n_grads) atomicOr(a.status, kErrGradient);
                    } else if (id >= (uint32_t)a.n_uses) {
                        atomicOr(a.status, kErrPicture);
                    }
                }
            }
            for (int o = o0; o < o1; ++o) {
                const int e0 = a.ooff[o], e1 = a.ooff[o + 1];
                const Runs rr = count_runs(a.pairs + 4 * (size_t)e0, e1 - e0);
                units += rr.units;
                strands += rr.strands;
            }
        }
        a.keep[s] = keep ? 1u : 0u;
        a.units[s] = units;
        a.strands[s] = strands;
        a.box[s] = make_float4(l, t, r, b);
    }
}

__device__ void emit_body(const Args &a, const Orders &ord, int t0, int ts)
{
    for (int s = t0; s < a.n_shapes; s += ts) {
        if (!a.keep[s]) continue;
        const uint64_t idx = a.kidx[s];
        uint64_t base = a.ustart[s];
        uint64_t cbase = a.curves ? a.cstart[s] : 0;
        const int o0 = a.soff[s], o1 = a.soff[s + 1];
        for (int o = o0; o < o1; ++o) {
            const int e0 = a.ooff[o], e1 = a.ooff[o + 1];
            const float *p = a.pairs + 4 * (size_t)e0;
            const Runs rr = count_runs(p, e1 - e0);
            emit_outline(p, e1 - e0, rr, a.heap, base, ord, a.curves, cbase, e0,
                         (a.tags[s] & dev::kTagNonzero) != 0);
            base += rr.units;
            cbase += rr.units - rr.strands;
        }
        ShapeRec rec;
        rec.tag = a.tags[s];
        rec.geo_start = (uint32_t)a.ustart[s];
        rec.num_strands = a.strands[s];
        a.shapes[idx] = rec;
        a.bbox[idx] = a.box[s];
    }
}

// Extent of a shape's strands: x from the end points, y over every point; a
// NaN y disables the y test (the host's record validation computes the same).
__device__ void extents_body(const Args &a, int n, int t0, int ts)
{
    for (int i = t0; i < n; i += ts) {
        const ShapeRec rec = a.shapes[i];
        size_t pos = 16u * (size_t)rec.geo_start;
        float xlo = INFINITY, xhi = -INFINITY, ylo = INFINITY, yhi = -INFINITY;
        bool y_nan = false;
        for (uint32_t k = 0; k < rec.num_strands; ++k) {
            const uint32_t size = *reinterpret_cast<const uint32_t *>(a.heap + pos) & 0xFFFFu;
            const float *pts = reinterpret_cast<const float *>(a.heap + pos + 8);
            const int npts = (int)size - 1;
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
        a.bounds[i] = make_float4(xlo, ylo, xhi, yhi);
    }
}

__global__ void __launch_bounds__(kThreads) ser_count(Args a) { count_body(a, gtid(), gstride()); }

__global__ void __launch_bounds__(kThreads) ser_emit(Args a, Orders ord) { emit_body(a, ord, gtid(), gstride()); }

__global__ void __launch_bounds__(kThreads) ser_extents(Args a, const uint64_t *n_kept)
{
    extents_body(a, (int)*n_kept, gtid(), gstride());
}

__global__ void ser_reset(uint32_t *status) { if (threadIdx.x < 4) status[threadIdx.x] = 0u; }

__device__ __forceinline__ int row_tid() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }
__device__ __forceinline__ int row_stride() { return (int)(gridDim.x * blockDim.x); }

__global__ void ser_reset_many(const Args *args, int n)
{
    for (int s = row_tid(); s < n; s += row_stride())
        for (int k = 0; k < 4; ++k) args[s].status[k] = 0u;
}

__global__ void __launch_bounds__(kThreads) ser_count_many(const Args *args)
{
    count_body(args[blockIdx.y], row_tid(), row_stride());
}

__global__ void __launch_bounds__(kThreads) ser_emit_many(const Args *args, Orders ord)
{
    emit_body(args[blockIdx.y], ord, row_tid(), row_stride());
}

__global__ void __launch_bounds__(kThreads) ser_extents_many(const Args *args)
{
    const Args &a = args[blockIdx.y];
    extents_body(a, (int)a.kidx[a.n_shapes], row_tid(), row_stride());
}

// Scan inputs.
struct KeepFlags {
    const uint32_t *keep;
    int n;
    __device__ void load() {}
    __device__ int count() const { return n; }
    __device__ uint64_t operator()(int i) const { return keep[i]; }
};
struct Units {
    const uint64_t *units;
    int n;
    __device__ void load() {}
    __device__ int count() const { return n; }
    __device__ uint64_t operator()(int i) const { return units[i]; }
};
// curves of a shape: each strand of m curves takes m + 1 units
struct Curves {
    const uint64_t *units;
    const uint32_t *strands;
    int n;
    __device__ void load() {}
    __device__ int count() const { return n; }
    __device__ uint64_t operator()(int i) const { return units[i] - strands[i]; }
};

// One block per scene: its three scans, then its totals and status.
__global__ void __launch_bounds__(kThreads) ser_scans_many(const Args *args, uint64_t *totals)
{
    const Args &a = args[blockIdx.x];
    const int n = a.n_shapes;
    devscan::block_range_excl_scan(KeepFlags{a.keep, n}, a.kidx);
    __syncthreads();
    devscan::block_range_excl_scan(Units{a.units, n}, a.ustart);
    __syncthreads();
    if (a.cstart) {
        devscan::block_range_excl_scan(Curves{a.units, a.strands, n}, a.cstart);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        uint64_t *t = totals + 4 * (size_t)blockIdx.x;
        t[0] = a.kidx[n];
        t[1] = a.ustart[n];
        t[2] = a.cstart ? a.cstart[n] : 0;
        t[3] = a.status[0];
    }
}

Orders make_orders()
{
    Orders o;
    std::memset(&o, 0, sizeof(o));
    for (int m = 1; m <= kMaxStrandCurves; ++m) {
        const int *ord = strand_tree_order(m);
        for (int k = 0; k + 1 < m; ++k) o.o[m][k] = (int8_t)ord[k];
    }
    return o;
}

int grid_for(int n)
{
    return std::max(1, std::min(4096, (n + kThreads - 1) / kThreads));
}

}  // namespace
}  // namespace ser
}  // namespace gudni

namespace gudni {

int check_outline_offsets(gudni_ctx *ctx, const gudni_outlines *o, int32_t &n_outlines, int32_t &n_pairs)
{
    const int32_t n = o->n_shapes;
    n_outlines = n ? o->shape_outline_offsets[n] : 0;
    if (n && o->shape_outline_offsets[0] < 0)
        return set_error(ctx, GUDNI_ERR_ARG, "shape outline offsets must be non-decreasing from 0");
    for (int32_t i = 0; i < n; ++i)
        if (o->shape_outline_offsets[i] < 0 || o->shape_outline_offsets[i] > o->shape_outline_offsets[i + 1])
            return set_error(ctx, GUDNI_ERR_ARG, "shape outline offsets must be non-decreasing from 0");
    n_pairs = n_outlines ? o->outline_pair_offsets[n_outlines] : 0;
    for (int32_t j = 0; j < n_outlines; ++j)
        if (o->outline_pair_offsets[j] < 0 || o->outline_pair_offsets[j] > o->outline_pair_offsets[j + 1])
            return set_error(ctx, GUDNI_ERR_ARG, "outline pair offsets must be non-decreasing from 0");
    if (n_pairs && !o->pairs) return set_error(ctx, GUDNI_ERR_ARG, "outlines without pairs");
    return GUDNI_OK;
}

int serialize_outlines(gudni_ctx *ctx, const SerIn &in, const SerOut &out, SerCounts &cnt)
{
    const hipStream_t st = ctx->stream;
    const int32_t n = in.n_shapes;
    CHECK_HIP(ctx, ctx->ser_work.ensure(16u * (size_t)n + 8u * (4 * (size_t)n + 3) + 4u * 2 * (size_t)n +
                                        8u * (devscan::kScanBlocks + 1) + 64));
    ser::Args a;
    std::memset(&a, 0, sizeof(a));
    uint64_t *partials;
    {
        // float4 first, then 8-byte and 4-byte arrays: every array aligned
        uint8_t *w = (uint8_t *)ctx->ser_work.p;
        a.box = (float4 *)w;              w += 16u * (size_t)n;
        a.units = (uint64_t *)w;          w += 8u * (size_t)n;
        a.kidx = (uint64_t *)w;           w += 8u * ((size_t)n + 1);
        a.ustart = (uint64_t *)w;         w += 8u * ((size_t)n + 1);
        a.cstart = (uint64_t *)w;         w += 8u * ((size_t)n + 1);
        a.keep = (uint32_t *)w;           w += 4u * (size_t)n;
        a.strands = (uint32_t *)w;        w += 4u * (size_t)n;
        a.status = (uint32_t *)w;         w += 16;
        w = (uint8_t *)(((uintptr_t)w + 15) & ~(uintptr_t)15);
        partials = (uint64_t *)w;
    }
    a.tags = in.tags;
    a.soff = in.soff;
    a.ooff = in.ooff;
    a.pairs = in.pairs;
    a.n_shapes = n;
    a.fw = (float)in.width;
    a.fh = (float)in.height;
    a.subs = in.subs;
    a.n_subs = in.n_subs;
    a.n_uses = in.n_uses;
    a.n_grads = in.n_grads;
    hipLaunchKernelGGL(ser::ser_reset, dim3(1), dim3(64), 0, st, a.status);
    hipLaunchKernelGGL(ser::ser_count, dim3(ser::grid_for(n)), dim3(ser::kThreads), 0, st, a);
    devscan::excl_scan(ser::KeepFlags{a.keep, n}, partials, a.kidx, st);
    devscan::excl_scan(ser::Units{a.units, n}, partials, a.ustart, st);
    if (out.curves) devscan::excl_scan(ser::Curves{a.units, a.strands, n}, partials, a.cstart, st);
    CHECK_HIP(ctx, hipGetLastError());
    uint64_t totals[3] = {0, 0, 0};
    uint32_t status = 0;
    CHECK_HIP(ctx, hipMemcpyAsync(&totals[0], a.kidx + n, 8, hipMemcpyDeviceToHost, st));
    CHECK_HIP(ctx, hipMemcpyAsync(&totals[1], a.ustart + n, 8, hipMemcpyDeviceToHost, st));
    if (out.curves) CHECK_HIP(ctx, hipMemcpyAsync(&totals[2], a.cstart + n, 8, hipMemcpyDeviceToHost, st));
    CHECK_HIP(ctx, hipMemcpyAsync(&status, a.status, 4, hipMemcpyDeviceToHost, st));
    CHECK_HIP(ctx, hipStreamSynchronize(st));
    if (status & ser::kErrSubstance)
        return set_error(ctx, GUDNI_ERR_ARG, "shape substance id outside the substance table");
    if (status & ser::kErrGradient)
        return set_error(ctx, GUDNI_ERR_ARG, "gradient substance refers to a missing gradient");
    if (status & ser::kErrPicture)
        return set_error(ctx, GUDNI_ERR_ARG, "picture substance refers to a missing usage");
    const uint64_t n_kept = totals[0], units = totals[1];
    if (units >= (1ull << 32) || totals[2] >= (1ull << 31))
        return set_error(ctx, GUDNI_ERR_ARG, "serialised geometry above 64 GiB or 2^31 curves");
    CHECK_HIP(ctx, out.geo->ensure(16u * (size_t)units));
    CHECK_HIP(ctx, out.shapes->ensure(16u * (size_t)n_kept));
    CHECK_HIP(ctx, out.bbox->ensure(16u * (size_t)n_kept));
    CHECK_HIP(ctx, out.bounds->ensure(16u * (size_t)n_kept));
    if (out.curves) CHECK_HIP(ctx, out.curves->ensure(sizeof(CurveSrc) * (size_t)totals[2]));
    a.heap = (uint8_t *)out.geo->p;
    a.shapes = (ShapeRec *)out.shapes->p;
    a.bbox = (float4 *)out.bbox->p;
    a.bounds = (float4 *)out.bounds->p;
    a.curves = out.curves ? (CurveSrc *)out.curves->p : nullptr;
    hipLaunchKernelGGL(ser::ser_emit, dim3(ser::grid_for(n)), dim3(ser::kThreads), 0, st, a, ser::make_orders());
    hipLaunchKernelGGL(ser::ser_extents, dim3(ser::grid_for((int)n_kept)), dim3(ser::kThreads), 0, st, a,
                       (const uint64_t *)(a.kidx + n));
    CHECK_HIP(ctx, hipGetLastError());
    cnt.n_kept = n_kept;
    cnt.geo_bytes = 16u * units;
    cnt.n_curves = totals[2];
    return GUDNI_OK;
}

int serialize_outlines_many(gudni_ctx *ctx, const SerIn *in, const SerOut *out, SerCounts *cnt, int m)
{
    if (m <= 0) return GUDNI_OK;
    const hipStream_t st = ctx->stream;
    // per scene: box 16n | units, kidx, ustart, cstart 8(n+1) | keep, strands 4n | status 16
    auto ws_bytes = [](size_t n) {
        return (16u * n + 8u * (4 * n + 3) + 4u * 2 * n + 16 + 255) & ~(size_t)255;
    };
    std::vector<size_t> off((size_t)m);
    size_t total = 0;
    int32_t most = 0;
    for (int s = 0; s < m; ++s) {
        off[s] = total;
        total += ws_bytes((size_t)in[s].n_shapes);
        most = std::max(most, in[s].n_shapes);
    }
    CHECK_HIP(ctx, ctx->ser_many_ws.ensure(total + 32u * (size_t)m));
    std::vector<ser::Args> args((size_t)m);
    for (int s = 0; s < m; ++s) {
        ser::Args &a = args[s];
        std::memset(&a, 0, sizeof(a));
        const size_t n = (size_t)in[s].n_shapes;
        uint8_t *w = (uint8_t *)ctx->ser_many_ws.p + off[s];
        a.box = (float4 *)w;              w += 16u * n;
        a.units = (uint64_t *)w;          w += 8u * n;
        a.kidx = (uint64_t *)w;           w += 8u * (n + 1);
        a.ustart = (uint64_t *)w;         w += 8u * (n + 1);
        a.cstart = (uint64_t *)w;         w += 8u * (n + 1);
        a.keep = (uint32_t *)w;           w += 4u * n;
        a.strands = (uint32_t *)w;        w += 4u * n;
        a.status = (uint32_t *)w;
        if (!out[s].curves) a.cstart = nullptr;   // no curve tracking: no curve scan
        a.tags = in[s].tags;
        a.soff = in[s].soff;
        a.ooff = in[s].ooff;
        a.pairs = in[s].pairs;
        a.n_shapes = in[s].n_shapes;
        a.fw = (float)in[s].width;
        a.fh = (float)in[s].height;
        a.subs = in[s].subs;
        a.n_subs = in[s].n_subs;
        a.n_uses = in[s].n_uses;
        a.n_grads = in[s].n_grads;
    }
    uint64_t *totals_dev = (uint64_t *)((uint8_t *)ctx->ser_many_ws.p + total);
    CHECK_HIP(ctx, upload(ctx, ctx->ser_many_args, args.data(), sizeof(ser::Args) * args.size()));
    const ser::Args *dargs = (const ser::Args *)ctx->ser_many_args.p;
    const int gx = std::min(64, ser::grid_for(most));
    hipLaunchKernelGGL(ser::ser_reset_many, dim3(1), dim3(ser::kThreads), 0, st, dargs, m);
    hipLaunchKernelGGL(ser::ser_count_many, dim3(gx, m), dim3(ser::kThreads), 0, st, dargs);
    hipLaunchKernelGGL(ser::ser_scans_many, dim3(m), dim3(ser::kThreads), 0, st, dargs, totals_dev);
    CHECK_HIP(ctx, hipGetLastError());
    std::vector<uint64_t> totals(4 * (size_t)m);
    CHECK_HIP(ctx, hipMemcpyAsync(totals.data(), totals_dev, 8 * totals.size(), hipMemcpyDeviceToHost, st));
    CHECK_HIP(ctx, hipStreamSynchronize(st));
    int32_t most_kept = 0;
    for (int s = 0; s < m; ++s) {
        const uint64_t *t = totals.data() + 4 * (size_t)s;
        const uint32_t status = (uint32_t)t[3];
        const std::string where = "scene " + std::to_string(s) + ": ";
        if (status & ser::kErrSubstance)
            return set_error(ctx, GUDNI_ERR_ARG, where + "shape substance id outside the substance table");
        if (status & ser::kErrGradient)
            return set_error(ctx, GUDNI_ERR_ARG, where + "gradient substance refers to a missing gradient");
        if (status & ser::kErrPicture)
            return set_error(ctx, GUDNI_ERR_ARG, where + "picture substance refers to a missing usage");
        if (t[1] >= (1ull << 32) || t[2] >= (1ull << 31))
            return set_error(ctx, GUDNI_ERR_ARG, where + "serialised geometry above 64 GiB or 2^31 curves");
        const SerOut &o = out[s];
        CHECK_HIP(ctx, o.geo->ensure(16u * (size_t)t[1]));
        CHECK_HIP(ctx, o.shapes->ensure(16u * (size_t)t[0]));
        CHECK_HIP(ctx, o.bbox->ensure(16u * (size_t)t[0]));
        CHECK_HIP(ctx, o.bounds->ensure(16u * (size_t)t[0]));
        if (o.curves) CHECK_HIP(ctx, o.curves->ensure(sizeof(CurveSrc) * (size_t)t[2]));
        ser::Args &a = args[s];
        a.heap = (uint8_t *)o.geo->p;
        a.shapes = (ShapeRec *)o.shapes->p;
        a.bbox = (float4 *)o.bbox->p;
        a.bounds = (float4 *)o.bounds->p;
        a.curves = o.curves ? (CurveSrc *)o.curves->p : nullptr;
        cnt[s].n_kept = t[0];
        cnt[s].geo_bytes = 16u * t[1];
        cnt[s].n_curves = t[2];
        most_kept = std::max(most_kept, (int32_t)t[0]);
    }
    CHECK_HIP(ctx, upload(ctx, ctx->ser_many_args, args.data(), sizeof(ser::Args) * args.size()));
    hipLaunchKernelGGL(ser::ser_emit_many, dim3(gx, m), dim3(ser::kThreads), 0, st, dargs, ser::make_orders());
    hipLaunchKernelGGL(ser::ser_extents_many, dim3(std::min(64, ser::grid_for(most_kept)), m), dim3(ser::kThreads),
                       0, st, dargs);
    CHECK_HIP(ctx, hipGetLastError());
    return GUDNI_OK;
}

}  // namespace gudni

using namespace gudni;

extern "C" {

int gudni_scene_upload_outlines(gudni_ctx *ctx, const gudni_outlines *o, const gudni_scene *sc)
{
    if (!ctx || !o || !sc) return GUDNI_ERR_ARG;
    if (sc->geo || sc->geo_bytes || sc->shapes || sc->n_shapes || sc->bbox)
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_scene_upload_outlines: the scene's geometry, shape "
                                             "and box fields must be empty (the device builds them)");
    if (o->n_shapes < 0 || sc->width <= 0 || sc->height <= 0 ||
        (o->n_shapes && (!o->tags || !o->shape_outline_offsets || !o->outline_pair_offsets)))
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_scene_upload_outlines: bad outlines or canvas");
    if (std::max(sc->width, sc->height) > (1 << 15))
        return set_error(ctx, GUDNI_ERR_ARG, "gudni_scene_upload_outlines: canvas side above 32768");
    const int32_t n = o->n_shapes;
    int32_t n_outlines = 0, n_pairs = 0;
    int rc = check_outline_offsets(ctx, o, n_outlines, n_pairs);
    if (rc) return rc;
    ctx->scene_ready = false;
    rc = gudni_upload_common(ctx, nullptr, 0, sc->substances, sc->n_substances, sc->pict_data,
                             sc->pict_bytes, sc->pict_uses, sc->n_pict_uses, sc->random_field,
                             sc->n_random);
    if (rc) return rc;
    rc = gudni_upload_gradients(ctx, sc->gradients, sc->n_gradients);
    if (rc) return rc;
    const hipStream_t st = ctx->stream;
    const bool timed = ctx->timing;
    if (timed)
        for (hipEvent_t &e : ctx->ser_ev)
            if (!e) CHECK_HIP(ctx, hipEventCreate(&e));
    CHECK_HIP(ctx, upload(ctx, ctx->ser_tags, o->tags, 8u * (size_t)n));
    CHECK_HIP(ctx, upload(ctx, ctx->ser_soff, o->shape_outline_offsets, 4u * ((size_t)n + 1)));
    CHECK_HIP(ctx, upload(ctx, ctx->ser_ooff, o->outline_pair_offsets, 4u * ((size_t)n_outlines + 1)));
    CHECK_HIP(ctx, upload(ctx, ctx->ser_pairs, o->pairs, 16u * (size_t)n_pairs));
    SerIn in;
    in.tags = (const uint64_t *)ctx->ser_tags.p;
    in.soff = (const int32_t *)ctx->ser_soff.p;
    in.ooff = (const int32_t *)ctx->ser_ooff.p;
    in.pairs = (const float *)ctx->ser_pairs.p;
    in.n_shapes = n;
    in.width = sc->width;
    in.height = sc->height;
    in.subs = (const float4 *)ctx->subs.p;
    in.n_subs = ctx->n_substances;
    in.n_uses = ctx->n_uses;
    in.n_grads = ctx->n_grads;
    SerOut out{&ctx->geo, &ctx->scene_shapes, &ctx->scene_bbox, &ctx->scene_bounds, nullptr};
    SerCounts cnt;
    if (timed) CHECK_HIP(ctx, hipEventRecord(ctx->ser_ev[0], st));
    rc = serialize_outlines(ctx, in, out, cnt);
    if (rc) return rc;
    if (timed) CHECK_HIP(ctx, hipEventRecord(ctx->ser_ev[1], st));
    CHECK_HIP(ctx, hipStreamSynchronize(st));
    if (timed) CHECK_HIP(ctx, hipEventElapsedTime(&ctx->last_ser_ms, ctx->ser_ev[0], ctx->ser_ev[1]));
    // the heap lives on the device only: host-side record checks of later
    // host-built jobs have nothing to check against
    ctx->geo_host.clear();
    ctx->scene_geo_bytes = (size_t)cnt.geo_bytes;
    ctx->scene_n_shapes = (int32_t)cnt.n_kept;
    ctx->scene_w = sc->width;
    ctx->scene_h = sc->height;
    std::memcpy(ctx->scene_bg, sc->background, sizeof(ctx->scene_bg));
    ctx->scene_ready = true;
    return GUDNI_OK;
}

int gudni_scene_geometry_sizes(gudni_ctx *ctx, int64_t sizes[2])
{
    if (!ctx || !sizes) return GUDNI_ERR_ARG;
    if (!ctx->scene_ready) return set_error(ctx, GUDNI_ERR_ARG, "no scene was uploaded");
    sizes[0] = ctx->scene_n_shapes;
    sizes[1] = (int64_t)ctx->scene_geo_bytes;
    return GUDNI_OK;
}

int gudni_scene_geometry_read(gudni_ctx *ctx, void *geo, void *shapes, float *bbox, float *bounds)
{
    if (!ctx) return GUDNI_ERR_ARG;
    if (!ctx->scene_ready) return set_error(ctx, GUDNI_ERR_ARG, "no scene was uploaded");
    CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t n = (size_t)ctx->scene_n_shapes;
    if (geo && ctx->scene_geo_bytes)
        CHECK_HIP(ctx, hipMemcpy(geo, ctx->geo.p, ctx->scene_geo_bytes, hipMemcpyDeviceToHost));
    if (shapes && n) CHECK_HIP(ctx, hipMemcpy(shapes, ctx->scene_shapes.p, 16 * n, hipMemcpyDeviceToHost));
    if (bbox && n) CHECK_HIP(ctx, hipMemcpy(bbox, ctx->scene_bbox.p, 16 * n, hipMemcpyDeviceToHost));
    if (bounds && n) CHECK_HIP(ctx, hipMemcpy(bounds, ctx->scene_bounds.p, 16 * n, hipMemcpyDeviceToHost));
    return GUDNI_OK;
}

float gudni_last_serialize_ms(const gudni_ctx *ctx) { return ctx ? ctx->last_ser_ms : 0.0f; }

}  // extern "C"
