// bin.h -- stage (a) on the device: the adaptive tile tree of
// buildTileTree / addShapeToTree (Raster/TileTree.hs:81-190) and the job
// layout of buildRasterJobs (Raster/Job.hs:132-179, OpenCL/CallKernels.hs:244-255),
// built level by level from the shapes' boxes.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "internal.h"

namespace gudni {
namespace binning {

constexpr int kBlocks = 512;    // grid of every grid-stride binning kernel
constexpr int kThreads = 256;

struct Params {
    int32_t n_shapes;
    int32_t canvas_depth;       // ceil(log2(max(W, H)))
    int32_t tile_depth;         // log2(max_tile_size)
    int32_t cap_shapes;         // checkTileSpace: shapeCount < maxShapes - 1
    uint32_t cap_strands;       // checkTileSpace: strand sum < maxStrandsPerTile
    int32_t per_job;            // tiles per raster job
    int32_t col_step;           // columnAllocation step
    int32_t threads_per_tile;
    int32_t width, height;
    int32_t max_levels;
};

// Device workspace; every array is carved from one allocation.
struct Workspace {
    int32_t cap_pairs = 0, cap_nodes = 0, max_levels = 0;
    int32_t *pair_shape = nullptr, *pair_node = nullptr;   // cap_pairs
    uint64_t *scan = nullptr;                              // cap_pairs + 1
    uint64_t *partials = nullptr;                          // kBlocks + 1
    int4 *node_box = nullptr, *node_meta = nullptr;        // cap_nodes
    int32_t *node_child = nullptr, *node_leaves = nullptr, *node_refs = nullptr;
    int32_t *node_toff = nullptr, *node_roff = nullptr;    // cap_nodes
    int4 *levels = nullptr;        // per level: node begin, node count, pair begin, pair count
    uint32_t *status = nullptr;    // [0] overflow [1] tiles [2] refs [3] work [4] nodes [5] pairs
};

size_t workspace_bytes(int32_t cap_pairs, int32_t cap_nodes, int32_t max_levels);
void carve(Workspace &w, void *base, int32_t cap_pairs, int32_t cap_nodes, int32_t max_levels);

// Outputs (capacities: tiles and tile_job cap_nodes, refs and ref_bounds
// cap_pairs, work 4*cap_nodes).
struct Output {
    TileRec *tiles;
    int32_t *tile_job;
    ShapeRec *refs;
    float4 *ref_bounds;
    uint32_t *work;
};

// Enqueues the whole binning on `stream`; status[] holds the counts once the
// stream reaches it (status[0] != 0: nothing valid; 1 = a capacity was
// exceeded, status[6] bit 0 = nodes, bit 1 = pairs; 2 = the tree is deeper
// than the level table).
hipError_t enqueue(const Params &p, const Workspace &w, const ShapeRec *shapes,
                   const float4 *bbox, const float4 *shape_bounds, const Output &out,
                   hipStream_t stream);

// One scene of a many-scene binning: its parameters, its own workspace and
// outputs.
struct SceneJob {
    Params p;
    Workspace w;
    const ShapeRec *shapes;
    const float4 *bbox, *shape_bounds;
    Output out;
};

constexpr int kSceneBlockThreads = 512;

// Bins n scenes in one launch, one block per scene: every level's steps and
// scans run in the block between barriers, so a small scene costs one
// block's sweep instead of a dozen launches per level.  `jobs` is in device
// memory; each job's status[] holds its counts once the stream reaches it.
hipError_t enqueue_scenes(const SceneJob *jobs, int n, hipStream_t stream);

}  // namespace binning
}  // namespace gudni
