// internal.h -- shared types of libgudni_hip's host and device code.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

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

struct Geometry {
    std::vector<uint8_t> heap;
    std::vector<ShapeRec> shapes;
    std::vector<BBox> bbox;
    std::vector<uint32_t> strands;
    std::vector<int32_t> source;
};

struct Jobs {
    std::vector<TileRec> tiles;
    std::vector<ShapeRec> refs;
    std::vector<int32_t> job_tile_offset;
    std::vector<int32_t> job_shape_offset;
};

inline int ceil_log2(int x)
{   // adjustedLog (TileTree.hs:74-75): ceiling (log2 x), 0 for x < 1
    int k = 0;
    if (x < 1) return 0;
    while ((1LL << k) < (long long)x) ++k;
    return k;
}

extern thread_local std::string g_host_error;

} // namespace gudni
