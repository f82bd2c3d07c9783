// serialize.h -- the device serialiser as a building block: outlines already
// in device memory -> strand heap, Shape records, boxes, strand extents and
// (optionally) per-curve provenance, into buffers the caller owns.
#pragma once
#include "context.h"

namespace gudni {

struct SerIn {
    const uint64_t *tags;          // device: per shape
    const int32_t *soff, *ooff;    // device: shape -> outlines, outline -> pairs
    const float *pairs;            // device: [n_pairs][4]
    int32_t n_shapes;
    int32_t width, height;
    const float4 *subs;            // device substance table (substance checks)
    int32_t n_subs, n_uses, n_grads;
};

struct SerOut {
    DevBuf *geo, *shapes, *bbox, *bounds;
    DevBuf *curves;                // null: no curve tracking
};

struct SerCounts {
    uint64_t n_kept = 0, geo_bytes = 0, n_curves = 0;
};

// Offsets of an onShape sequence: non-decreasing from 0 and inside their
// arrays (the kernels index with them).  Sets the outline and pair counts.
int check_outline_offsets(gudni_ctx *ctx, const gudni_outlines *o, int32_t &n_outlines, int32_t &n_pairs);

// Stream-ordered on ctx->stream with one synchronisation (the sizes); grows
// the output buffers as needed.  Offsets must already be validated.
int serialize_outlines(gudni_ctx *ctx, const SerIn &in, const SerOut &out, SerCounts &cnt);

// The same for m scenes in a handful of launches (a grid row per scene, a
// block per scene for the scans) and one synchronisation in all.
int serialize_outlines_many(gudni_ctx *ctx, const SerIn *in, const SerOut *out, SerCounts *cnt, int m);

}  // namespace gudni
