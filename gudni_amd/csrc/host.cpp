// host.cpp -- host stages of the raster path for libgudni_hip: scene ->
// strand heap + shape records (the Haskell onShape serialiser) and the
// adaptive tile tree flattened into raster jobs (the Haskell tile tree and
// job builder).  Everything here runs once per frame on the CPU and produces
// the exact byte layouts the device kernels (and the reference's kernels)
// consume.
#include "gudni_hip.h"
#include "internal.h"

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace gudni {

thread_local std::string g_host_error;

// --------------------------------------------------------------------------
// Strand tree order.
// ReorderTable.hs:44-111 lays the interior (point, control) pairs of a strand
// out as the breadth-first order of a left-complete binary search tree.  That
// is the Eytzinger layout: array slot k has children 2k+1 and 2k+2, and an
// in-order walk of the slots visits the pairs in increasing order.
// --------------------------------------------------------------------------
static void eytzinger_fill(int n, int slot, int &next, int *pair_of_slot)
{
    if (slot >= n) return;
    eytzinger_fill(n, 2 * slot + 1, next, pair_of_slot);
    pair_of_slot[slot] = next++;
    eytzinger_fill(n, 2 * slot + 2, next, pair_of_slot);
}

struct TreeOrders {
    // order[m] = pair index of each slot, for a strand of m curves (m-1 nodes)
    int order[kMaxStrandCurves + 1][kMaxStrandCurves];
    TreeOrders()
    {
        for (int m = 1; m <= kMaxStrandCurves; ++m) {
            int next = 0;
            eytzinger_fill(m - 1, 0, next, order[m]);
        }
    }
};
static const TreeOrders g_orders;

const int *strand_tree_order(int m) { return g_orders.order[m]; }

struct Pt { float x, y; };
struct Quad { Pt p0, c, p1; };

static inline Pt lerp_t(float t, Pt a, Pt b)
{
    // between t v0 v1 = v0 ^* (1-t) ^+^ v1 ^* t   (Deknob.hs:40-41)
    const float s = 1.0f - t;
    return Pt{a.x * s + b.x * t, a.y * s + b.y * t};
}

// Split a quadratic whose control point lies strictly outside the x-range of
// its end points ("knob") at the parameter where dx/dt changes sign, found by
// the bisection of Deknob.hs:61-79 on FP32 parameters.
static int split_knob(const Quad &q, Quad *out, float *split_t)
{
    const bool bulge_left = q.c.x < q.p0.x && q.c.x < q.p1.x;
    const bool bulge_right = q.c.x > q.p0.x && q.c.x > q.p1.x;
    *split_t = 0.0f;
    if (!bulge_left && !bulge_right) {
        out[0] = q;
        return 1;
    }
    auto outward = [bulge_left](Pt a, Pt b) { return bulge_left ? a.x < b.x : a.x > b.x; };
    float lo = 0.0f, hi = 1.0f, t = 0.5f;
    Pt m0, on, m1;
    for (;;) {
        m0 = lerp_t(t, q.p0, q.c);
        m1 = lerp_t(t, q.c, q.p1);
        on = lerp_t(t, m0, m1);
        if (hi - lo <= 0.0001f) break;           // iota (Figure/Space.hs:101-102)
        if (outward(m1, on)) {                   // still bulging: move toward p1
            const float nt = t + ((hi - t) / 2.0f);
            lo = t;
            t = nt;
        } else if (outward(m0, on)) {            // overshot: move toward p0
            const float nt = lo + ((t - lo) / 2.0f);
            hi = t;
            t = nt;
        } else {
            break;
        }
    }
    out[0] = Quad{q.p0, m0, on};
    out[1] = Quad{on, m1, q.p1};
    *split_t = t;
    return 2;
}

static inline int x_direction(const Quad &q)
{
    return (q.p0.x < q.p1.x) ? -1 : (q.p0.x > q.p1.x ? 1 : 0);
}

// Append one oriented strand of m <= 16 curves (Strand.hs:180-206):
// header {u16 size in 8-byte units, u16 0, u32 0}, rightmost point, leftmost
// point + its control, then the interior (point, control) pairs in tree order.
// flags: the header's second u16 (GUDNI_STRAND_REVERSED for a strand of a
// nonzero-fill shape that runs against its outline; 0 otherwise).
static void emit_strand(const Quad *qs, int m, std::vector<uint8_t> &heap, uint32_t curve_base = 0,
                        uint16_t flags = 0)
{
    const size_t base = heap.size();
    const size_t bytes = 16u * (size_t)(m + 1);
    heap.resize(base + bytes);
    uint8_t *w = heap.data() + base;
    const uint16_t size = (uint16_t)(bytes / 8);
    std::memset(w, 0, 8);
    std::memcpy(w, &size, 2);
    std::memcpy(w + 2, &flags, 2);
    std::memcpy(w + 4, &curve_base, 4);   // zero unless curves are tracked
    float *f = reinterpret_cast<float *>(w + 8);
    f[0] = qs[m - 1].p1.x; f[1] = qs[m - 1].p1.y;
    f[2] = qs[0].p0.x;     f[3] = qs[0].p0.y;
    f[4] = qs[0].c.x;      f[5] = qs[0].c.y;
    const int *order = g_orders.order[m];
    for (int k = 0; k < m - 1; ++k) {
        const Quad &node = qs[1 + order[k]];
        f[6 + 4 * k + 0] = node.p0.x; f[6 + 4 * k + 1] = node.p0.y;
        f[6 + 4 * k + 2] = node.c.x;  f[6 + 4 * k + 3] = node.c.y;
    }
}

// One closed outline -> strands (Strand.hs:153-178).  Returns strand count.
// curves (optional): records each strand curve's origin; pair_base is the
// global index of the outline's first pair.
static int outline_strands(const float *pairs, int k, std::vector<Quad> &work,
                           std::vector<uint8_t> &heap, std::vector<CurveSrc> *curves = nullptr,
                           int pair_base = 0, bool mark_reversed = false)
{
    if (k < 2) return 0;
    work.clear();
    std::vector<CurveSrc> wsrc;
    for (int i = 0; i < k; ++i) {
        const int j = (i + 1 == k) ? 0 : i + 1;
        Quad q{{pairs[4 * i], pairs[4 * i + 1]}, {pairs[4 * i + 2], pairs[4 * i + 3]},
               {pairs[4 * j], pairs[4 * j + 1]}};
        Quad two[2];
        float t = 0.0f;
        const int n = split_knob(q, two, &t);
        work.push_back(two[0]);
        if (n == 2) work.push_back(two[1]);
        if (curves) {
            const int32_t a = 2 * (pair_base + i), b = 2 * (pair_base + j);
            wsrc.push_back(CurveSrc{a, a + 1, b, t, n == 2 ? kCurveLowerHalf : 0u});
            if (n == 2) wsrc.push_back(CurveSrc{a, a + 1, b, t, kCurveUpperHalf});
        }
    }
    // Maximal runs of curves that share a non-zero x direction; the run that
    // is still open when the loop ends is emitted first (Strand.hs:77-102).
    std::vector<int> cuts;
    cuts.push_back(0);
    for (size_t i = 1; i < work.size(); ++i) {
        const int a = x_direction(work[i - 1]), b = x_direction(work[i]);
        if (a != b || a == 0) cuts.push_back((int)i);
    }
    cuts.push_back((int)work.size());
    const int runs = (int)cuts.size() - 1;
    int strands = 0;
    Quad oriented[kMaxStrandCurves];
    for (int r = 0; r < runs; ++r) {
        const int run = (r == 0) ? runs - 1 : r - 1;
        for (int s = cuts[run]; s < cuts[run + 1]; s += kMaxStrandCurves) {
            const int m = std::min(kMaxStrandCurves, cuts[run + 1] - s);
            const Quad *src = &work[s];
            const bool reversed = src[0].p0.x > src[m - 1].p1.x;
            if (reversed) {   // run right-to-left: reverse
                for (int q = 0; q < m; ++q) {
                    const Quad &o = src[m - 1 - q];
                    oriented[q] = Quad{o.p1, o.c, o.p0};
                }
            } else {
                std::copy(src, src + m, oriented);
            }
            const uint16_t flags = (mark_reversed && reversed) ? GUDNI_STRAND_REVERSED : 0;
            if (curves) {
                emit_strand(oriented, m, heap, (uint32_t)curves->size(), flags);
                for (int q = 0; q < m; ++q) {
                    CurveSrc c = wsrc[(size_t)s + (size_t)(reversed ? m - 1 - q : q)];
                    if (reversed) c.flags |= kCurveReversed;
                    curves->push_back(c);
                }
            } else {
                emit_strand(oriented, m, heap, 0, flags);
            }
            ++strands;
        }
    }
    return strands;
}

int build_geometry(const gudni_outlines *sc, int32_t cw, int32_t ch, Geometry &g, bool track_curves)
{
    g.heap.clear(); g.shapes.clear(); g.bbox.clear(); g.strands.clear(); g.source.clear();
    g.curves.clear();
    const float fw = (float)cw, fh = (float)ch;
    std::vector<Quad> work;
    for (int32_t i = 0; i < sc->n_shapes; ++i) {
        const int o0 = sc->shape_outline_offsets[i], o1 = sc->shape_outline_offsets[i + 1];
        const int p0 = sc->outline_pair_offsets[o0], p1 = sc->outline_pair_offsets[o1];
        if (p1 <= p0) continue;
        // bounding box over every on- and off-curve point (Outline.hs:123-133)
        float l = sc->pairs[4 * p0], t = sc->pairs[4 * p0 + 1], r = l, b = t;
        for (int p = p0; p < p1; ++p) {
            const float *q = sc->pairs + 4 * (size_t)p;
            for (int h = 0; h < 2; ++h) {
                const float x = q[2 * h], y = q[2 * h + 1];
                if (!(l <= x)) l = x;
                if (!(x <= r)) r = x;
                if (!(t <= y)) t = y;
                if (!(y <= b)) b = y;
            }
        }
        if (l >= fw || t >= fh || r <= 0.0f || b <= 0.0f) continue;   // excludeBox
        const size_t start = g.heap.size();
        uint32_t n = 0;
        for (int o = o0; o < o1; ++o) {
            const int a = sc->outline_pair_offsets[o], e = sc->outline_pair_offsets[o + 1];
            n += (uint32_t)outline_strands(sc->pairs + 4 * (size_t)a, e - a, work, g.heap,
                                           track_curves ? &g.curves : nullptr, a,
                                           (sc->tags[i] & GUDNI_TAG_NONZERO) != 0);
        }
        ShapeRec rec;
        rec.tag = sc->tags[i];
        rec.geo_start = (uint32_t)(start / 16);
        rec.num_strands = n;
        g.shapes.push_back(rec);
        g.bbox.push_back(BBox{l, t, r, b});
        g.strands.push_back(n);
        g.source.push_back(i);
    }
    return GUDNI_OK;
}

// --------------------------------------------------------------------------
// Tile binning.
// TileTree.hs:113-190 inserts shapes one by one and splits a leaf when it
// overflows.  Capacity is monotone in the set of shapes a node receives and
// routing depends only on each bounding box, so the final tree is fixed by
// the sets alone: a node below the tile size is a leaf iff its whole set fits
// (<= maxShapes-1 shapes and fewer than maxStrandsPerTile strands) or it is at
// the 8-pixel floor; its list is its set in reverse scene order.  This builds
// that tree directly by recursive partition.
// --------------------------------------------------------------------------
struct Binner {
    const Geometry &g;
    int cap_shapes;
    uint64_t cap_strands;
    int tile_depth;
    std::vector<TileRec> tiles;
    std::vector<int> lists;      // leaf lists, concatenated, newest first
    std::vector<int> list_start;

    Binner(const Geometry &geo, const gudni_spec &s) : g(geo)
    {
        cap_shapes = s.max_shapes - 1;
        cap_strands = (uint64_t)s.max_strands_per_tile;
        tile_depth = ceil_log2(s.max_tile_size);
    }

    void leaf(const int *box, int hd, int vd, const std::vector<int> &set)
    {
        TileRec t;
        t.box[0] = box[0]; t.box[1] = box[1]; t.box[2] = box[2]; t.box[3] = box[3];
        t.hdepth = (int16_t)hd; t.vdepth = (int16_t)vd;
        t.column_allocation = 0; t.shape_start = 0; t.shape_count = (int32_t)set.size();
        list_start.push_back((int)lists.size());
        for (size_t i = set.size(); i-- > 0;) lists.push_back(set[i]);
        tiles.push_back(t);
    }

    bool fits(const std::vector<int> &set) const
    {
        if ((int)set.size() > cap_shapes) return false;
        uint64_t s = 0;
        for (int i : set) s += g.strands[i];
        return s < cap_strands;
    }

    // vertical == true: a VTree/VLeaf node (split across a horizontal cut)
    void node(bool vertical, const int *box, int hd, int vd, const std::vector<int> &set,
              bool forced)
    {
        if (!forced) {
            const bool floor = vertical ? (box[3] - box[1]) <= 8 : (box[2] - box[0]) <= 8;
            if (floor || fits(set)) {
                leaf(box, hd, vd, set);
                return;
            }
        }
        std::vector<int> a, b;
        if (vertical) {
            const int cut = box[1] + (box[3] - box[1]) / 2;
            const float fc = (float)cut;
            for (int i : set) {
                if (g.bbox[i].t < fc) a.push_back(i);
                if (g.bbox[i].b > fc) b.push_back(i);
            }
            const int ta[4] = {box[0], box[1], box[2], cut};
            const int tb[4] = {box[0], cut, box[2], box[3]};
            // forced (root) splits keep both depths; adaptive V splits make
            // 2:1 HLeafs of depth (hd, vd-1)
            if (forced) {
                root_h(hd, ta, a);
                root_h(hd, tb, b);
            } else {
                node(false, ta, hd, vd - 1, a, false);
                node(false, tb, hd, vd - 1, b, false);
            }
        } else {
            const int cut = box[0] + (box[2] - box[0]) / 2;
            const float fc = (float)cut;
            for (int i : set) {
                if (g.bbox[i].l < fc) a.push_back(i);
                if (g.bbox[i].r > fc) b.push_back(i);
            }
            const int la[4] = {box[0], box[1], cut, box[3]};
            const int lb[4] = {cut, box[1], box[2], box[3]};
            node(true, la, hd - 1, vd, a, false);
            node(true, lb, hd - 1, vd, b, false);
        }
    }

    // buildTileTree's goV / goH (TileTree.hs:81-109): square pre-split down to
    // the tile depth; the pre-split leaves are VLeafs of depth (d, d).
    void root_v(int depth, const int *box, const std::vector<int> &set)
    {
        if (depth > tile_depth) node(true, box, depth, depth, set, true);
        else node(true, box, depth, depth, set, false);
    }
    void root_h(int depth, const int *box, const std::vector<int> &set)
    {
        const int cut = box[0] + (1 << (depth - 1));
        const float fc = (float)cut;
        std::vector<int> a, b;
        for (int i : set) {
            if (g.bbox[i].l < fc) a.push_back(i);
            if (g.bbox[i].r > fc) b.push_back(i);
        }
        const int la[4] = {box[0], box[1], cut, box[3]};
        const int lb[4] = {cut, box[1], box[2], box[3]};
        root_v(depth - 1, la, a);
        root_v(depth - 1, lb, b);
    }
};

int build_jobs(const gudni_spec &spec, const Geometry &g, int32_t cw, int32_t ch, Jobs &out)
{
    Binner bn(g, spec);
    out.sort_radix_bits = spec.sort_radix_bits;
    const int canvas_depth = ceil_log2(std::max(cw, ch));
    const int side = 1 << canvas_depth;
    const int root[4] = {0, 0, side, side};
    std::vector<int> all((size_t)g.shapes.size());
    for (size_t i = 0; i < all.size(); ++i) all[i] = (int)i;
    bn.root_v(canvas_depth, root, all);

    // Job.hs:151-179 with the argument order of CallKernels.hs:254
    // (maxTilesPerJob <- threadsPerTile, column step <- tilesPerCall).
    const int per_job = spec.threads_per_tile;
    const int col_step = spec.max_tiles_per_call;
    out.tiles.clear(); out.refs.clear(); out.job_tile_offset.assign(1, 0);
    out.job_shape_offset.assign(1, 0);
    int in_job = 0, col = 0;
    for (size_t t = 0; t < bn.tiles.size(); ++t) {
        if (in_job >= per_job) {
            out.job_tile_offset.push_back((int32_t)out.tiles.size());
            out.job_shape_offset.push_back((int32_t)out.refs.size());
            in_job = 0;
            col = 0;
        }
        TileRec rec = bn.tiles[t];
        rec.shape_start = (int32_t)out.refs.size() - out.job_shape_offset.back();
        rec.column_allocation = col;
        const int s0 = bn.list_start[t];
        for (int k = 0; k < rec.shape_count; ++k) out.refs.push_back(g.shapes[bn.lists[s0 + k]]);
        out.tiles.push_back(rec);
        ++in_job;
        col += col_step;
    }
    out.job_tile_offset.push_back((int32_t)out.tiles.size());
    out.job_shape_offset.push_back((int32_t)out.refs.size());
    return GUDNI_OK;
}

} // namespace gudni

using namespace gudni;

struct gudni_geometry { Geometry g; };

extern "C" {

void gudni_default_spec(gudni_spec *s)
{
    s->max_tile_size = 256;
    s->threads_per_tile = 256;
    s->max_tiles_per_call = 256;
    s->max_thresholds = 4096;
    s->max_strands_per_tile = 4094;
    s->max_shapes = 127;
    s->random_field_size = 4096;
    s->stochastic_factor = 0.0f;
    s->coverage_mode = GUDNI_COVERAGE_ACCURATE;
    s->texture_filter = GUDNI_FILTER_NEAREST;
    s->sort_radix_bits = 4;
}

int gudni_geometry_build(const gudni_outlines *scene, int32_t cw, int32_t ch, gudni_geometry **out)
{
    if (!scene || !out || cw <= 0 || ch <= 0 || scene->n_shapes < 0) {
        g_host_error = "gudni_geometry_build: bad argument";
        return GUDNI_ERR_ARG;
    }
    try {
        auto *g = new gudni_geometry();
        int rc = build_geometry(scene, cw, ch, g->g, false);
        if (rc) { delete g; return rc; }
        *out = g;
        return GUDNI_OK;
    } catch (const std::bad_alloc &) {
        g_host_error = "gudni_geometry_build: out of memory";
        return GUDNI_ERR_ARG;
    }
}

int gudni_geometry_view_get(const gudni_geometry *g, gudni_geometry_view *v)
{
    if (!g || !v) return GUDNI_ERR_ARG;
    v->geo = g->g.heap.data();
    v->geo_bytes = g->g.heap.size();
    v->shapes = reinterpret_cast<const uint8_t *>(g->g.shapes.data());
    v->bbox = reinterpret_cast<const float *>(g->g.bbox.data());
    v->strand_count = g->g.strands.data();
    v->source_index = g->g.source.data();
    v->n_shapes = (int32_t)g->g.shapes.size();
    return GUDNI_OK;
}

void gudni_geometry_free(gudni_geometry *g) { delete g; }

int gudni_jobs_build(const gudni_spec *spec, const gudni_geometry *g, int32_t cw, int32_t ch,
                     gudni_jobs **out)
{
    if (!spec || !g || !out || cw <= 0 || ch <= 0) {
        g_host_error = "gudni_jobs_build: bad argument";
        return GUDNI_ERR_ARG;
    }
    if (spec->max_tile_size <= 0 || (spec->max_tile_size & (spec->max_tile_size - 1)) ||
        spec->max_tile_size > spec->threads_per_tile ||
        spec->threads_per_tile <= 0 || spec->max_tiles_per_call <= 0 ||
        spec->sort_radix_bits < 0 || spec->sort_radix_bits > 5) {
        g_host_error = "gudni_jobs_build: bad spec";
        return GUDNI_ERR_ARG;
    }
    try {
        auto *j = new gudni_jobs();
        int rc = build_jobs(*spec, g->g, cw, ch, j->j);
        if (rc) { delete j; return rc; }
        *out = j;
        return GUDNI_OK;
    } catch (const std::bad_alloc &) {
        g_host_error = "gudni_jobs_build: out of memory";
        return GUDNI_ERR_ARG;
    }
}

int gudni_jobs_view_get(const gudni_jobs *j, gudni_jobs_view *v)
{
    if (!j || !v) return GUDNI_ERR_ARG;
    v->n_jobs = (int32_t)j->j.job_tile_offset.size() - 1;
    v->job_tile_offset = j->j.job_tile_offset.data();
    v->job_shape_offset = j->j.job_shape_offset.data();
    v->tiles = reinterpret_cast<const uint8_t *>(j->j.tiles.data());
    v->shape_refs = reinterpret_cast<const uint8_t *>(j->j.refs.data());
    v->n_tiles = (int32_t)j->j.tiles.size();
    v->n_refs = (int32_t)j->j.refs.size();
    return GUDNI_OK;
}

void gudni_jobs_free(gudni_jobs *j) { delete j; }

} // extern "C"
