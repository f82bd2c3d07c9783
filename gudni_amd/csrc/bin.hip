// bin.hip -- stage (a) on the device: shapes binned into the adaptive tile
// tree and laid out as raster jobs.
//
// The tree of buildTileTree / addShapeToTree (Raster/TileTree.hs:81-190) is
// fixed by the set of shapes each node receives: a node is a leaf when it is
// at the 8-pixel floor or its shapes fit (count < maxShapes - 1 and strand sum
// < maxStrandsPerTile, checkTileSpace :167-172), otherwise it splits at its
// midline and a shape goes to each child its box strictly crosses into
// (insertShapeV/H :117-151).  The root is pre-split in alternating halves down
// to the tile size (buildTileTree :81-109).  So the tree is built here one
// level at a time from (node, shape) pairs kept in scene order:
//
//   decide    one thread per node: leaf or split
//   scans     exclusive scans of the split flags (child numbering) and of the
//             per-pair child flags (child pair ranges)
//   children  split nodes write their two children
//   scatter   pairs move to the children they route to, order kept
//
// then leaf counts go up the tree, DFS offsets come down it (traverseTileTree
// :193-204: top before bottom, left before right), and every leaf writes its
// TileInfo and its shape references newest first (:158) at its DFS position,
// with the job fields of accumulateRasterJobs (Raster/Job.hs:151-179,
// arguments as CallKernels.hs:254 passes them) and the reversed job numbering
// of queueRasterJobs (CallKernels.hs:241, 255).  Every kernel is grid-stride
// over counts it reads from device memory, so the whole build is enqueued
// without a host round trip; capacities are checked on the device.
#include "bin.h"
#include "scan.h"

namespace gudni {
namespace binning {
namespace {

constexpr uint32_t kFlagVertical = 1u, kFlagForced = 2u, kFlagSplit = 4u;

__device__ __forceinline__ uint32_t lo32(uint64_t v) { return (uint32_t)v; }
__device__ __forceinline__ uint32_t hi32(uint64_t v) { return (uint32_t)(v >> 32); }

__device__ __forceinline__ int node_cut(const int4 b, bool vertical)
{
    return vertical ? b.y + (b.w - b.y) / 2 : b.x + (b.z - b.x) / 2;
}

// (child a, child b) membership of a shape box: strict tests against the cut
__device__ __forceinline__ uint64_t route(const float4 bb, const int4 box, bool vertical)
{
    const float fc = (float)node_cut(box, vertical);
    const bool fa = vertical ? bb.y < fc : bb.x < fc;
    const bool fb = vertical ? bb.w > fc : bb.z > fc;
    return (uint64_t)(fa ? 1u : 0u) | ((uint64_t)(fb ? 1u : 0u) << 32);
}

// Scan inputs: a count read from device memory and a value per index.
struct SplitFlags {
    const int4 *levels, *meta;
    int level;
    int begin, n;
    __device__ void load() { const int4 lv = levels[level]; begin = lv.x; n = lv.y; }
    __device__ int count() const { return n; }
    __device__ uint64_t operator()(int i) const
    {
        return (meta[begin + i].y & kFlagSplit) ? 1u : 0u;
    }
};

struct PairFlags {
    const int4 *levels, *meta, *box;
    const int32_t *pair_shape, *pair_node;
    const float4 *bbox;
    int level;
    int begin, n;
    __device__ void load() { const int4 lv = levels[level]; begin = lv.z; n = lv.w; }
    __device__ int count() const { return n; }
    __device__ uint64_t operator()(int i) const
    {
        const int p = begin + i;
        const int node = pair_node[p];
        const uint32_t flags = (uint32_t)meta[node].y;
        if (!(flags & kFlagSplit)) return 0;
        return route(bbox[pair_shape[p]], box[node], (flags & kFlagVertical) != 0);
    }
};

// isActiveThread (Kernels.cl:1720-1722) over the columns of work unit qd
__device__ bool quarter_active(const TileRec &r, int qd, int cd, int width, int height)
{
    const int hd = r.hdepth, vd = r.vdepth;
    const int dd = max(0, vd - (cd - hd));
    for (int l = 0; l < 64; ++l) {
        const int col = unit_column(hd, cd, qd, l);
        const int ix = ((1 << hd) - 1) & col, iy = (col >> hd) << dd;
        if (iy < (r.box[3] - r.box[1]) && ix + r.box[0] < width && iy + r.box[1] < height)
            return true;
    }
    return false;
}

struct WorkFlags {
    const TileRec *tiles;
    const uint32_t *status;
    int cd, quarters, width, height;
    int n;
    __device__ void load() { n = status[0] ? 0 : (int)status[1]; }
    __device__ int count() const { return n; }
    __device__ uint64_t operator()(int t) const
    {
        const TileRec r = tiles[t];
        uint64_t k = 0;
        for (int q = 0; q < quarters; ++q) k += quarter_active(r, q, cd, width, height) ? 1u : 0u;
        return k;
    }
};

// ---------------------------------------------------------------------------
// tree levels

struct Dev {
    Params p;
    int32_t *pair_shape, *pair_node;
    uint64_t *scan, *node_scan;
    int4 *node_box, *node_meta;
    int32_t *node_child, *node_leaves, *node_refs, *node_toff, *node_roff;
    int4 *levels;
    uint32_t *status;
    int32_t cap_pairs, cap_nodes;
};

__device__ __forceinline__ int ceil_log2_dev(int x)
{
    int k = 0;
    while ((1 << k) < x) ++k;
    return k;
}

__device__ __forceinline__ int gtid() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }
__device__ __forceinline__ int gstride() { return (int)(gridDim.x * blockDim.x); }

// Each step is a device function over (first thread, stride): the grid-stride
// kernels below call them once per launch, and the one-block-per-scene kernel
// calls them all in turn with block barriers in between.
__device__ void init_body(const Dev &d, int t0, int ts)
{
    for (int i = t0; i < d.p.n_shapes; i += ts) {
        d.pair_shape[i] = i;
        d.pair_node[i] = 0;
    }
    for (int l = t0; l < d.p.max_levels; l += ts)
        d.levels[l] = l == 0 ? make_int4(0, 1, 0, d.p.n_shapes) : make_int4(0, 0, 0, 0);
    if (t0 == 0) {
        const int side = 1 << d.p.canvas_depth;
        const int cd = d.p.canvas_depth;
        d.node_box[0] = make_int4(0, 0, side, side);
        const uint32_t flags = kFlagVertical | (cd > d.p.tile_depth ? kFlagForced : 0u);
        d.node_meta[0] = make_int4((cd & 0xFFFF) | (cd << 16), (int)flags, 0, d.p.n_shapes);
        d.node_toff[0] = 0;
        d.node_roff[0] = 0;
        for (int k = 0; k < 8; ++k) d.status[k] = 0;
    }
}

__global__ void __launch_bounds__(kThreads) bin_init(Dev d) { init_body(d, gtid(), gstride()); }

__device__ void decide_body(const Dev &d, const ShapeRec *shapes, int level, int t0, int ts)
{
    if (d.status[0]) return;
    const int4 lv = d.levels[level];
    for (int i = t0; i < lv.y; i += ts) {
        const int n = lv.x + i;
        const int4 m = d.node_meta[n];
        uint32_t flags = (uint32_t)m.y;
        bool split;
        if (flags & kFlagForced) {
            split = true;
        } else {
            const int4 b = d.node_box[n];
            const bool floor = (flags & kFlagVertical) ? (b.w - b.y) <= 8 : (b.z - b.x) <= 8;
            if (floor) {
                split = false;
            } else if (m.w > d.p.cap_shapes) {
                split = true;
            } else {
                uint64_t s = 0;
                for (int k = 0; k < m.w; ++k) s += shapes[d.pair_shape[m.z + k]].num_strands;
                split = !(s < (uint64_t)d.p.cap_strands);
            }
        }
        if (split) flags |= kFlagSplit;
        d.node_meta[n].y = (int)flags;
        d.node_child[n] = -1;
    }
}

__global__ void __launch_bounds__(kThreads) bin_decide(Dev d, const ShapeRec *shapes, int level)
{
    decide_body(d, shapes, level, gtid(), gstride());
}

// Next level's extent from the two scans; checks the capacities.
__device__ void advance_body(const Dev &d, int level)
{
    if (d.status[0]) return;
    const int4 lv = d.levels[level];
    const int nb = lv.x + lv.y, pb = lv.z + lv.w;
    const uint64_t ps = d.scan[lv.w];
    const int nc = 2 * (int)d.node_scan[lv.y];
    const int pc = (int)(lo32(ps) + hi32(ps));
    if (level + 1 >= d.p.max_levels) {
        if (nc > 0) d.status[0] = 2;
        return;
    }
    const bool over_nodes = (int64_t)nb + nc > d.cap_nodes, over_pairs = (int64_t)pb + pc > d.cap_pairs;
    if (over_nodes || over_pairs) {
        // which capacity ran out: bit 0 nodes, bit 1 (node, shape) pairs
        d.status[6] = (over_nodes ? 1u : 0u) | (over_pairs ? 2u : 0u);
        d.status[0] = 1;
        return;
    }
    d.levels[level + 1] = make_int4(nb, nc, pb, pc);
}

__global__ void bin_advance(Dev d, int level) { advance_body(d, level); }

__device__ void children_body(const Dev &d, int level, int t0, int ts)
{
    if (d.status[0]) return;
    const int4 lv = d.levels[level];
    const int nb = lv.x + lv.y, pb = lv.z + lv.w;
    for (int i = t0; i < lv.y; i += ts) {
        const int n = lv.x + i;
        const int4 m = d.node_meta[n];
        const uint32_t flags = (uint32_t)m.y;
        if (!(flags & kFlagSplit)) continue;
        const int a = nb + 2 * (int)d.node_scan[i];
        const int s = m.z - lv.z, e = s + m.w;
        const uint64_t ss = d.scan[s], se = d.scan[e];
        const int a_start = pb + (int)(lo32(ss) + hi32(ss));
        const int a_count = (int)(lo32(se) - lo32(ss));
        const int b_count = (int)(hi32(se) - hi32(ss));
        const int4 box = d.node_box[n];
        const bool vertical = (flags & kFlagVertical) != 0;
        const bool forced = (flags & kFlagForced) != 0;
        const int hd = m.x & 0xFFFF, vd = m.x >> 16;
        const int cut = node_cut(box, vertical);
        int4 ba, bb;
        int chd, cvd;
        uint32_t cflags;
        if (vertical) {
            ba = make_int4(box.x, box.y, box.z, cut);
            bb = make_int4(box.x, cut, box.z, box.w);
            // forced V (root pre-split): H children of the same depth;
            // adaptive V: 2:1 HLeaf children of depth (hd, vd - 1)
            chd = hd;
            cvd = forced ? vd : vd - 1;
            cflags = forced ? kFlagForced : 0u;
        } else {
            ba = make_int4(box.x, box.y, cut, box.w);
            bb = make_int4(cut, box.y, box.z, box.w);
            if (forced) {
                chd = cvd = hd - 1;
                cflags = kFlagVertical | (hd - 1 > d.p.tile_depth ? kFlagForced : 0u);
            } else {
                chd = hd - 1;
                cvd = vd;
                cflags = kFlagVertical;
            }
        }
        const int dm = (chd & 0xFFFF) | (cvd << 16);
        d.node_box[a] = ba;
        d.node_box[a + 1] = bb;
        d.node_meta[a] = make_int4(dm, (int)cflags, a_start, a_count);
        d.node_meta[a + 1] = make_int4(dm, (int)cflags, a_start + a_count, b_count);
        d.node_child[n] = a;
    }
}

__global__ void __launch_bounds__(kThreads) bin_children(Dev d, int level) { children_body(d, level, gtid(), gstride()); }

__device__ void scatter_body(const Dev &d, const float4 *bbox, int level, int t0, int ts)
{
    if (d.status[0]) return;
    const int4 lv = d.levels[level];
    for (int i = t0; i < lv.w; i += ts) {
        const int p = lv.z + i;
        const int n = d.pair_node[p];
        const int4 m = d.node_meta[n];
        const uint32_t flags = (uint32_t)m.y;
        if (!(flags & kFlagSplit)) continue;
        const int shape = d.pair_shape[p];
        const uint64_t f = route(bbox[shape], d.node_box[n], (flags & kFlagVertical) != 0);
        const int c = d.node_child[n];
        const uint64_t ss = d.scan[m.z - lv.z], si = d.scan[i];
        if (lo32(f)) {
            const int pos = d.node_meta[c].z + (int)(lo32(si) - lo32(ss));
            d.pair_shape[pos] = shape;
            d.pair_node[pos] = c;
        }
        if (hi32(f)) {
            const int pos = d.node_meta[c + 1].z + (int)(hi32(si) - hi32(ss));
            d.pair_shape[pos] = shape;
            d.pair_node[pos] = c + 1;
        }
    }
}

__global__ void __launch_bounds__(kThreads) bin_scatter(Dev d, const float4 *bbox, int level)
{
    scatter_body(d, bbox, level, gtid(), gstride());
}

__device__ void count_up_body(const Dev &d, int level, int t0, int ts)
{
    if (d.status[0]) return;
    const int4 lv = d.levels[level];
    for (int i = t0; i < lv.y; i += ts) {
        const int n = lv.x + i;
        const int c = d.node_child[n];
        if (c < 0) {
            d.node_leaves[n] = 1;
            d.node_refs[n] = d.node_meta[n].w;
        } else {
            d.node_leaves[n] = d.node_leaves[c] + d.node_leaves[c + 1];
            d.node_refs[n] = d.node_refs[c] + d.node_refs[c + 1];
        }
    }
}

__global__ void __launch_bounds__(kThreads) bin_count_up(Dev d, int level) { count_up_body(d, level, gtid(), gstride()); }

__device__ void offsets_down_body(const Dev &d, int level, int t0, int ts)
{
    if (d.status[0]) return;
    const int4 lv = d.levels[level];
    for (int i = t0; i < lv.y; i += ts) {
        const int n = lv.x + i;
        const int c = d.node_child[n];
        if (c < 0) continue;
        d.node_toff[c] = d.node_toff[n];
        d.node_toff[c + 1] = d.node_toff[n] + d.node_leaves[c];
        d.node_roff[c] = d.node_roff[n];
        d.node_roff[c + 1] = d.node_roff[n] + d.node_refs[c];
    }
}

__global__ void __launch_bounds__(kThreads) bin_offsets_down(Dev d, int level)
{
    offsets_down_body(d, level, gtid(), gstride());
}

__device__ void finish_body(const Dev &d)
{
    if (d.status[0]) return;
    int tn = 0, tp = 0;
    for (int l = 0; l < d.p.max_levels; ++l) {
        const int4 lv = d.levels[l];
        tn = max(tn, lv.x + lv.y);
        tp = max(tp, lv.z + lv.w);
    }
    d.status[1] = (uint32_t)d.node_leaves[0];
    d.status[2] = (uint32_t)d.node_refs[0];
    d.status[4] = (uint32_t)tn;
    d.status[5] = (uint32_t)tp;
}

__global__ void bin_finish(Dev d) { finish_body(d); }

__device__ void emit_tiles_body(const Dev &d, TileRec *tiles, int32_t *tile_job, int t0, int ts)
{
    if (d.status[0]) return;
    const int tn = (int)d.status[4];
    const int n_tiles = (int)d.status[1];
    const int n_jobs = (n_tiles + d.p.per_job - 1) / d.p.per_job;
    for (int n = t0; n < tn; n += ts) {
        if (d.node_child[n] >= 0) continue;
        const int t = d.node_toff[n];
        const int4 b = d.node_box[n];
        const int4 m = d.node_meta[n];
        TileRec r;
        r.box[0] = b.x; r.box[1] = b.y; r.box[2] = b.z; r.box[3] = b.w;
        r.hdepth = (int16_t)(m.x & 0xFFFF);
        r.vdepth = (int16_t)(m.x >> 16);
        r.column_allocation = (t % d.p.per_job) * d.p.col_step;
        r.shape_start = d.node_roff[n];
        r.shape_count = m.w;
        tiles[t] = r;
        tile_job[t] = n_jobs - 1 - t / d.p.per_job;
    }
}

__global__ void __launch_bounds__(kThreads) bin_emit_tiles(Dev d, TileRec *tiles, int32_t *tile_job)
{
    emit_tiles_body(d, tiles, tile_job, gtid(), gstride());
}

__device__ void emit_refs_body(const Dev &d, const ShapeRec *shapes, const float4 *shape_bounds,
                               ShapeRec *refs, float4 *ref_bounds, int t0, int ts)
{
    if (d.status[0]) return;
    const int tp = (int)d.status[5];
    for (int p = t0; p < tp; p += ts) {
        const int n = d.pair_node[p];
        if (d.node_child[n] >= 0) continue;
        const int4 m = d.node_meta[n];
        const int r = d.node_roff[n] + m.w - 1 - (p - m.z);   // newest first
        const int s = d.pair_shape[p];
        refs[r] = shapes[s];
        ref_bounds[r] = shape_bounds[s];
    }
}

__global__ void __launch_bounds__(kThreads) bin_emit_refs(Dev d, const ShapeRec *shapes,
                                                          const float4 *shape_bounds, ShapeRec *refs,
                                                          float4 *ref_bounds)
{
    emit_refs_body(d, shapes, shape_bounds, refs, ref_bounds, gtid(), gstride());
}

__device__ void emit_work_body(const Dev &d, WorkFlags f, uint32_t *work, int t0, int ts)
{
    if (d.status[0]) return;
    f.load();
    const int n = f.count();
    for (int t = t0; t < n; t += ts) {
        const TileRec r = f.tiles[t];
        uint32_t k = (uint32_t)d.scan[t];
        for (int q = 0; q < f.quarters; ++q)
            if (quarter_active(r, q, f.cd, f.width, f.height)) work[k++] = ((uint32_t)t << 2) | (uint32_t)q;
    }
    if (t0 == 0) d.status[3] = (uint32_t)d.scan[n];
}

__global__ void __launch_bounds__(kThreads) bin_emit_work(Dev d, WorkFlags f, uint32_t *work)
{
    emit_work_body(d, f, work, gtid(), gstride());
}

// One scene per block: the steps of enqueue() in order, with block barriers
// for the kernel boundaries and block scans for the device-wide ones.
__global__ void __launch_bounds__(kSceneBlockThreads) bin_scene_block(const SceneJob *jobs)
{
    const SceneJob &j = jobs[blockIdx.x];
    Dev d;
    d.p = j.p;
    d.pair_shape = j.w.pair_shape;
    d.pair_node = j.w.pair_node;
    d.scan = j.w.scan;
    d.node_scan = j.w.scan + (((size_t)8 * ((size_t)j.w.cap_pairs + 1) + 255) & ~(size_t)255) / 8;
    d.node_box = j.w.node_box;
    d.node_meta = j.w.node_meta;
    d.node_child = j.w.node_child;
    d.node_leaves = j.w.node_leaves;
    d.node_refs = j.w.node_refs;
    d.node_toff = j.w.node_toff;
    d.node_roff = j.w.node_roff;
    d.levels = j.w.levels;
    d.status = j.w.status;
    d.cap_pairs = j.w.cap_pairs;
    d.cap_nodes = j.w.cap_nodes;
    const int t0 = (int)threadIdx.x, ts = (int)blockDim.x;
    init_body(d, t0, ts);
    __syncthreads();
    int depth = d.p.max_levels;
    for (int l = 0; l < d.p.max_levels; ++l) {
        decide_body(d, j.shapes, l, t0, ts);
        __syncthreads();
        devscan::block_range_excl_scan(SplitFlags{d.levels, d.node_meta, l, 0, 0}, d.node_scan);
        __syncthreads();
        devscan::block_range_excl_scan(
            PairFlags{d.levels, d.node_meta, d.node_box, d.pair_shape, d.pair_node, j.bbox, l, 0, 0}, d.scan);
        __syncthreads();
        if (t0 == 0) advance_body(d, l);
        __syncthreads();
        children_body(d, l, t0, ts);
        __syncthreads();
        scatter_body(d, j.bbox, l, t0, ts);
        __syncthreads();
        // the levels below a level without nodes are empty: their steps
        // would do nothing (the same for every thread of the block)
        if (d.status[0] || l + 1 >= d.p.max_levels || d.levels[l + 1].y == 0) {
            depth = l + 1;
            break;
        }
    }
    for (int l = depth - 1; l >= 0; --l) {
        count_up_body(d, l, t0, ts);
        __syncthreads();
    }
    for (int l = 0; l < depth; ++l) {
        offsets_down_body(d, l, t0, ts);
        __syncthreads();
    }
    if (t0 == 0) finish_body(d);
    __syncthreads();
    emit_tiles_body(d, j.out.tiles, j.out.tile_job, t0, ts);
    emit_refs_body(d, j.shapes, j.shape_bounds, j.out.refs, j.out.ref_bounds, t0, ts);
    __syncthreads();
    WorkFlags wf{j.out.tiles, d.status, ceil_log2_dev(d.p.threads_per_tile), d.p.threads_per_tile / 64,
                 d.p.width, d.p.height, 0};
    devscan::block_range_excl_scan(wf, d.scan);
    __syncthreads();
    emit_work_body(d, wf, j.out.work, t0, ts);
}

}  // namespace

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t workspace_bytes(int32_t cap_pairs, int32_t cap_nodes, int32_t max_levels)
{
    size_t b = 0;
    b += 2 * align256(4 * (size_t)cap_pairs);
    b += align256(8 * ((size_t)cap_pairs + 1));
    b += align256(8 * ((size_t)cap_nodes + 1));
    b += align256(8 * ((size_t)kBlocks + 1));
    b += 2 * align256(16 * (size_t)cap_nodes);
    b += 5 * align256(4 * (size_t)cap_nodes);
    b += align256(16 * (size_t)max_levels);
    b += align256(4 * 8);
    return b;
}

// node_scan lives right after scan; carve() keeps the order of workspace_bytes()
static uint64_t *g_node_scan_of(const Workspace &w)
{
    return w.scan + align256(8 * ((size_t)w.cap_pairs + 1)) / 8;
}

void carve(Workspace &w, void *base, int32_t cap_pairs, int32_t cap_nodes, int32_t max_levels)
{
    uint8_t *p = (uint8_t *)base;
    auto take = [&p](size_t bytes) { void *r = p; p += align256(bytes); return r; };
    w.cap_pairs = cap_pairs;
    w.cap_nodes = cap_nodes;
    w.max_levels = max_levels;
    w.pair_shape = (int32_t *)take(4 * (size_t)cap_pairs);
    w.pair_node = (int32_t *)take(4 * (size_t)cap_pairs);
    w.scan = (uint64_t *)take(8 * ((size_t)cap_pairs + 1));
    (void)take(8 * ((size_t)cap_nodes + 1));   // node scan, see g_node_scan_of
    w.partials = (uint64_t *)take(8 * ((size_t)kBlocks + 1));
    w.node_box = (int4 *)take(16 * (size_t)cap_nodes);
    w.node_meta = (int4 *)take(16 * (size_t)cap_nodes);
    w.node_child = (int32_t *)take(4 * (size_t)cap_nodes);
    w.node_leaves = (int32_t *)take(4 * (size_t)cap_nodes);
    w.node_refs = (int32_t *)take(4 * (size_t)cap_nodes);
    w.node_toff = (int32_t *)take(4 * (size_t)cap_nodes);
    w.node_roff = (int32_t *)take(4 * (size_t)cap_nodes);
    w.levels = (int4 *)take(16 * (size_t)max_levels);
    w.status = (uint32_t *)take(4 * 8);
}

hipError_t enqueue(const Params &p, const Workspace &w, const ShapeRec *shapes, const float4 *bbox,
                   const float4 *shape_bounds, const Output &out, hipStream_t stream)
{
    Dev d;
    d.p = p;
    d.pair_shape = w.pair_shape;
    d.pair_node = w.pair_node;
    d.scan = w.scan;
    d.node_scan = g_node_scan_of(w);
    d.node_box = w.node_box;
    d.node_meta = w.node_meta;
    d.node_child = w.node_child;
    d.node_leaves = w.node_leaves;
    d.node_refs = w.node_refs;
    d.node_toff = w.node_toff;
    d.node_roff = w.node_roff;
    d.levels = w.levels;
    d.status = w.status;
    d.cap_pairs = w.cap_pairs;
    d.cap_nodes = w.cap_nodes;
    const dim3 g(kBlocks), b(kThreads);
    hipLaunchKernelGGL(bin_init, g, b, 0, stream, d);
    for (int l = 0; l < p.max_levels; ++l) {
        hipLaunchKernelGGL(bin_decide, g, b, 0, stream, d, shapes, l);
        SplitFlags sf{w.levels, w.node_meta, l, 0, 0};
        devscan::excl_scan(sf, w.partials, d.node_scan, stream);
        PairFlags pf{w.levels, w.node_meta, w.node_box, w.pair_shape, w.pair_node, bbox, l, 0, 0};
        devscan::excl_scan(pf, w.partials, w.scan, stream);
        hipLaunchKernelGGL(bin_advance, dim3(1), dim3(1), 0, stream, d, l);
        hipLaunchKernelGGL(bin_children, g, b, 0, stream, d, l);
        hipLaunchKernelGGL(bin_scatter, g, b, 0, stream, d, bbox, l);
    }
    for (int l = p.max_levels - 1; l >= 0; --l)
        hipLaunchKernelGGL(bin_count_up, g, b, 0, stream, d, l);
    for (int l = 0; l < p.max_levels; ++l)
        hipLaunchKernelGGL(bin_offsets_down, g, b, 0, stream, d, l);
    hipLaunchKernelGGL(bin_finish, dim3(1), dim3(1), 0, stream, d);
    hipLaunchKernelGGL(bin_emit_tiles, g, b, 0, stream, d, out.tiles, out.tile_job);
    hipLaunchKernelGGL(bin_emit_refs, g, b, 0, stream, d, shapes, shape_bounds, out.refs,
                       out.ref_bounds);
    WorkFlags wf{out.tiles, w.status, ceil_log2(p.threads_per_tile), p.threads_per_tile / 64,
                 p.width, p.height, 0};
    devscan::excl_scan(wf, w.partials, w.scan, stream);
    hipLaunchKernelGGL(bin_emit_work, g, b, 0, stream, d, wf, out.work);
    return hipGetLastError();
}

hipError_t enqueue_scenes(const SceneJob *jobs, int n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(bin_scene_block, dim3(n), dim3(kSceneBlockThreads), 0, stream, jobs);
    return hipGetLastError();
}

}  // namespace binning
}  // namespace gudni
