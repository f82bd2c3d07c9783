// internal.h -- shared types of libgudni_hip's host and device code.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "gudni_hip.h"

namespace gudni {

constexpr int kMaxStrandCurves = 16;   // splitTooLarge (Strand.hs:171-178)

// Shape record, 16 bytes (Raster/Types.hs:159-173, Kernels.cl:318-324)
struct ShapeRec {
    uint64_t tag;
    uint32_t geo_start;     // in 16-byte units of the strand heap
    uint32_t num_strands;
};
static_assert(sizeof(ShapeRec) == 16, "Shape record is 16 bytes");

// TileInfo, 32 bytes (Raster/Types.hs:176-198, Kernels.cl:335-341)
struct TileRec {
    int32_t box[4];         // left top right bottom
    int16_t hdepth, vdepth;
    int32_t column_allocation;
    int32_t shape_start;
    int32_t shape_count;
};
static_assert(sizeof(TileRec) == 32, "TileInfo is 32 bytes");

// PictureUse, 24 bytes (Figure/Picture.hs:183-199, Kernels.cl:350-355)
struct PictUse {
    float tx, ty;
    int32_t w, h;
    int32_t mem_offset;
    float scale;
};
static_assert(sizeof(PictUse) == 24, "PictureUse is 24 bytes");

struct BBox { float l, t, r, b; };

// Work units.  A tile of width 2^hd is run by 2^cd threads (cd = 8), one per
// (column, row segment) as initTileState numbers them (Kernels.cl:1692-1722:
// thread `col` takes x = col & (2^hd - 1) and segment col >> hd), in four
// units of 64 (one wave each).  A unit is a block of 2^cw columns by 2^sw
// segments, as square as the tile allows (8 x 8 where the tile has 8 or more
// segments): the x span of a unit is the filter of the shapes its wave walks,
// so a narrow block walks fewer.  Every column is computed the same whichever
// unit and lane runs it.  Unit q is the (q mod n_x)-th block across and the
// (q div n_x)-th down.
struct UnitShape {
    int cw_bits, sw_bits, n_x;
};
constexpr UnitShape unit_shape(int hd, int cd)
{
    int sw = cd - hd < 3 ? cd - hd : 3;
    if (sw < 0) sw = 0;
    int cw = 6 - sw;
    if (cw > hd) {
        cw = hd;
        sw = 6 - hd;
    }
    return UnitShape{cw, sw, 1 << (hd - cw)};
}
// The thread (reference column index) lane `lane` of unit q runs.
constexpr int unit_column(int hd, int cd, int q, int lane)
{
    const UnitShape u = unit_shape(hd, cd);
    const int x = ((q % u.n_x) << u.cw_bits) | (lane & ((1 << u.cw_bits) - 1));
    const int seg = ((q / u.n_x) << u.sw_bits) | (lane >> u.cw_bits);
    return (seg << hd) | x;
}
// First column (tile-relative) and width of unit q's block.
constexpr int unit_x0(int hd, int cd, int q) { return (q % unit_shape(hd, cd).n_x) << unit_shape(hd, cd).cw_bits; }
constexpr int unit_width(int hd, int cd) { return 1 << unit_shape(hd, cd).cw_bits; }

// Where a strand curve comes from, for gradients with respect to the outline
// points: the outline curve's on-curve start, control and on-curve end (point
// indices: pair q's on-curve point is 2q, its control 2q+1), the knob-split
// parameter and which half this is, and whether the strand runs it backwards.
struct CurveSrc {
    int32_t p0, c, p1;
    float t;
    uint32_t flags;
};
constexpr uint32_t kCurveLowerHalf = 1u;   // [0, t] of a knob split
constexpr uint32_t kCurveUpperHalf = 2u;   // [t, 1]
constexpr uint32_t kCurveReversed = 4u;    // strand order is p1 -> p0
static_assert(sizeof(CurveSrc) == 20, "CurveSrc is 20 bytes");

struct Geometry {
    std::vector<uint8_t> heap;
    std::vector<ShapeRec> shapes;
    std::vector<BBox> bbox;
    std::vector<uint32_t> strands;
    std::vector<int32_t> source;
    // with curve tracking only: one entry per strand curve, in heap order;
    // each strand header's second dword holds its first curve's index
    std::vector<CurveSrc> curves;
};

struct Jobs {
    int32_t sort_radix_bits = 0;   // of the spec the jobs were built with
    std::vector<TileRec> tiles;
    std::vector<ShapeRec> refs;
    std::vector<int32_t> job_tile_offset;
    std::vector<int32_t> job_shape_offset;
};

// Slot -> interior pair index of a strand of m curves (m - 1 tree slots):
// the Eytzinger order of ReorderTable.hs:44-111.
const int *strand_tree_order(int m);

inline int ceil_log2(int x)
{   // adjustedLog (TileTree.hs:74-75): ceiling (log2 x), 0 for x < 1
    int k = 0;
    if (x < 1) return 0;
    while ((1LL << k) < (long long)x) ++k;
    return k;
}

extern thread_local std::string g_host_error;

// Serialisation (track_curves: also record each strand curve's origin and
// write strand curve indices into the strand headers' second dword).
int build_geometry(const gudni_outlines *sc, int32_t cw, int32_t ch, Geometry &g,
                   bool track_curves);
int build_jobs(const gudni_spec &spec, const Geometry &g, int32_t cw, int32_t ch, Jobs &out);

} // namespace gudni

// The opaque handle of gudni_jobs_build: the jobs with the spec's sort radix.
struct gudni_jobs { gudni::Jobs j; };
