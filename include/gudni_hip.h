/*
 * gudni_hip.h -- C ABI of libgudni_hip, the MI355X (gfx950) raster path of
 * Gudni.  Plain C99: opaque handles, plain pointers and sizes, int status
 * codes (0 = ok, < 0 = error, message in gudni_last_error), no exceptions
 * across the boundary.  One gudni_ctx per thread; a ctx is not thread-safe.
 *
 * Record formats are the reference's own Storable layouts, so Haskell encoders
 * feed this library unchanged:
 *   strand heap    Raster/Strand.hs:180-206   (16*(m+1) bytes per strand)
 *   Shape          Raster/Types.hs:159-173    (u64 tag, u32 geoStart in
 *                                              16-byte units, u32 numStrands)
 *   TileInfo       Raster/Types.hs:176-198    (int4 box, i16 hDepth,
 *                                              i16 vDepth, i32 columnAllocation,
 *                                              i32 shapeStart, i32 shapeCount)
 *   Substance      Raster/ShapeInfo.hs:136-148 (float4 colour | u32 pictId)
 *   PictureUse     Figure/Picture.hs:183-199  (float2 translate, int2 size,
 *                                              i32 memOffset, f32 scale)
 * Output pixels are the reference's BGRA8 words (Kernels.cl:842-844).
 */
#ifndef GUDNI_HIP_H
#define GUDNI_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GUDNI_OK              0
#define GUDNI_ERR_ARG        -1   /* bad argument / size                      */
#define GUDNI_ERR_HIP        -2   /* HIP runtime error                        */
#define GUDNI_ERR_OVERFLOW   -3   /* a column's threshold queue exceeded
                                     maxThresholds (Kernels.cl:375-380 would
                                     corrupt the neighbouring column)         */
#define GUDNI_ERR_GUARD      -4   /* a column exceeded the section guard      */
#define GUDNI_ERR_NO_DEVICE  -5

/* RasterSpec (OpenCL/Rasterizer.hs:37-51) with the values determineRasterSpec
 * (OpenCL/Setup.hs:71-87) yields for a device reporting a maximum work-group
 * size of 256 and a 4 GiB maximum allocation. */
typedef struct gudni_spec {
    int32_t max_tile_size;        /* specMaxTileSize      = 256 (a power of
                                     two, at most threads_per_tile; 8..256
                                     for a context)                          */
    int32_t threads_per_tile;     /* specThreadsPerTile   = 256               */
    int32_t max_tiles_per_call;   /* specMaxTilesPerCall  = 256               */
    int32_t max_thresholds;       /* specMaxThresholds    = 4096              */
    int32_t max_strands_per_tile; /* specMaxStrandsPerTile = 4094             */
    int32_t max_shapes;           /* specMaxShapes        = 127               */
    int32_t random_field_size;    /* RANDOMFIELDSIZE      = 4096              */
    float   stochastic_factor;    /* STOCHASTIC_FACTOR    = 0.0               */
    int32_t coverage_mode;        /* GUDNI_COVERAGE_ACCURATE (default) or
                                     GUDNI_COVERAGE_REFERENCE                 */
    int32_t texture_filter;       /* GUDNI_FILTER_NEAREST (default: the
                                     reference's sampling, Kernels.cl:856-860)
                                     or GUDNI_FILTER_BILINEAR                 */
    int32_t sort_radix_bits;      /* 0..5: digit width of the radix pass that
                                     precedes the per-column insertion sort
                                     (sortThresholds, Kernels.cl:2084-2115);
                                     0 = insertion sort only.  The permutation
                                     is the same for every value.  An
                                     extension: the reference has no radix
                                     pass.                                    */
} gudni_spec;

/* Fill rule (an extension; the reference fills every shape by crossing
 * parity, Kernels.cl:298-304, 1515-1519).  A shape whose tag has this bit
 * set is filled where its winding number is non-zero; substance ids are the
 * tag's bits 58..0.  For such shapes the serialisers record, in bit 0 of a
 * strand header's second u16 (zero in the reference's layout,
 * Strand.hs:196-206), that the strand runs against its outline
 * (reverseIfBackwards, Strand.hs:137-143 reversed it); strands of even-odd
 * shapes keep the reference's bytes. */
#define GUDNI_TAG_NONZERO      0x0800000000000000ull
#define GUDNI_STRAND_REVERSED  0x0001u

/* Picture sampling at the pixel's integer position p (Kernels.cl:1426):
 * s = p / scale - translate.  NEAREST takes the texel at rtz(s); BILINEAR
 * blends the four texels around s - 0.5 (texels outside the picture are
 * transparent). */
#define GUDNI_FILTER_NEAREST  0
#define GUDNI_FILTER_BILINEAR 1

/* Gradient substances (no counterpart in the reference): a shape whose tag
 * has both substance-type bits set (0xC000...) takes its colour from the
 * gradient record its substance's first word indexes, evaluated at the
 * pixel's integer position like a picture.  112 bytes:
 *   kind 0: linear, t = ((p - p0) . (p1 - p0)) / |p1 - p0|^2
 *   kind 1: radial, t = |p - p0| / p1.x
 * t clamped to [0, 1]; the colour interpolates the straight RGBA stops
 * linearly between offsets (offsets ascending, 1..4 stops). */
typedef struct gudni_gradient {
    int32_t kind;
    int32_t n_stops;
    float   x0, y0, x1, y1;
    float   unused[2];
    float   offset[4];
    float   rgba[4][4];
} gudni_gradient;
#define GUDNI_GRADIENT_LINEAR 0
#define GUDNI_GRADIENT_RADIAL 1

/* Evaluation of section areas and colours.  Records, their order, the
 * section structure and the shapes each colour walk visits are identical in
 * both modes.
 * Section mid-x (thresholdMidXLow, Kernels.cl:916-926):
 * REFERENCE evaluates yMid = y0 + (y1 - y0)/2 and then x(yMid) exactly as the
 * reference writes it; in tall column segments yMid's rounding (ulp(200) =
 * 1.5e-5) times a shallow edge's inverse slope costs up to ~1e-5 of
 * coverage.  ACCURATE evaluates the same point as
 * ((y0 - start.y) + (y1 - y0)/2) * invSlope + left, whose intermediates are
 * bounded by the threshold's own height (coverage within ~1e-6 of an FP64
 * integration of the same sections).
 * Compositing (determineColor, Kernels.cl:1447-1513): REFERENCE
 * un-premultiplies after every shape (three divisions per step, composite,
 * Kernels.cl:878-887); ACCURATE keeps the accumulated colour premultiplied
 * (P += rgb*a*(1-A), A += a*(1-A) with the reference's alpha expression, so
 * the walk stops at the same shape) and divides once per walk; on a tile of
 * n shapes, all added with distinct substances, it composites four groups of
 * ceil(n/4) consecutive shapes and then the groups (the same "over", rounded
 * as that grouping rounds it).  Frames of the two modes differ by at most one
 * 8-bit step. */
#define GUDNI_COVERAGE_ACCURATE  0
#define GUDNI_COVERAGE_REFERENCE 1

typedef struct gudni_ctx gudni_ctx;

/* Fills the pinned default spec above. */
void        gudni_default_spec(gudni_spec *spec);
/* Replaces setupOpenCL (OpenCL/Setup.hs:102-147): selects HIP device
 * `device`, records the spec; kernels are compiled ahead of time. */
int         gudni_create(const gudni_spec *spec, int device, gudni_ctx **out);
void        gudni_destroy(gudni_ctx *ctx);
const char *gudni_last_error(const gudni_ctx *ctx);
/* Launch stream (a hipStream_t); NULL = the null stream. */
int         gudni_set_stream(gudni_ctx *ctx, void *hip_stream);

/* ---- Level 1: byte-compatible with generateCall -------------------------- */
/* Replaces the common-buffer upload in queueRasterJobs
 * (OpenCL/CallKernels.hs:224-240): geometry pile, substance pile, picture
 * data, picture usages and the random field.  Copies to device memory. */
int gudni_upload_common(gudni_ctx *ctx,
                        const void *geo, size_t geo_bytes,
                        const void *substances, int32_t n_substances,
                        const void *pict_data, size_t pict_bytes,
                        const void *pict_uses, int32_t n_pict_uses,
                        const float *random_field, int32_t n_random);

/* Gradient records (gudni_gradient) for the scene of gudni_upload_common;
 * an extension with no counterpart in the reference. */
int  gudni_upload_gradients(gudni_ctx *ctx, const void *gradients, int32_t n_gradients);

/* Replaces generateCall (OpenCL/CallKernels.hs:88-179): runs the
 * generateThresholds / sortThresholds / renderThresholds work of one
 * RasterJob (Raster/Job.hs:72-77) and writes the job's pixels into `out`
 * (host memory, width*height words; other pixels are left untouched: only
 * the rows the job's columns reach are copied to the device and back -- a
 * column writes 2^dd rows from its origin, clipped to the canvas but not to
 * its tile's box, Kernels.cl:1692-1722).
 * Synchronous, like the reference's KernelSync launches. */
int gudni_raster_job(gudni_ctx *ctx,
                     const void *shapes, int32_t n_shapes,
                     const void *tiles, int32_t n_tiles,
                     const float background[4],
                     int32_t width, int32_t height,
                     int32_t frame, int32_t job_index,
                     uint32_t *out);

/* ---- Host scene preparation (the Haskell host stages) -------------------- */
/* A scene as the sequence of onShape calls buildOverScene makes
 * (Raster/Serialize.hs:148-177, 266-273): per shape its ShapeTag and its
 * transformed outlines, each a closed loop of curve pairs
 * (Figure/Outline.hs:55-107; straight segments carry their midpoint). */
typedef struct gudni_outlines {
    int32_t         n_shapes;
    const uint64_t *tags;                  /* [n_shapes]                      */
    const int32_t  *shape_outline_offsets; /* [n_shapes+1]                    */
    const int32_t  *outline_pair_offsets;  /* [n_outlines+1]                  */
    const float    *pairs;                 /* [n_pairs][4] on.x on.y off.x off.y */
} gudni_outlines;

/* Serialised geometry: what onShape leaves in the geometry pile plus the
 * ShapeEntry fields (Raster/Types.hs:140-145) addShapeToTree reads. */
typedef struct gudni_geometry_view {
    const uint8_t  *geo;      size_t geo_bytes;
    const uint8_t  *shapes;   /* 16-byte Shape records, kept shapes only      */
    const float    *bbox;     /* [n][4] left top right bottom                 */
    const uint32_t *strand_count;
    const int32_t  *source_index;          /* index in gudni_outlines         */
    int32_t         n_shapes;
} gudni_geometry_view;

typedef struct gudni_geometry gudni_geometry;
/* Replaces onShape's serialisation (Serialize.hs:148-177, Enclosure.hs:62-73,
 * Strand.hs:153-206): culls off-canvas shapes (excludeBox, Serialize.hs:97-104)
 * and writes strands. */
int  gudni_geometry_build(const gudni_outlines *scene, int32_t canvas_w,
                          int32_t canvas_h, gudni_geometry **out);
int  gudni_geometry_view_get(const gudni_geometry *g, gudni_geometry_view *v);
void gudni_geometry_free(gudni_geometry *g);

/* Raster jobs: the tile tree (Raster/TileTree.hs:81-204) flattened into jobs
 * of <= max_tiles_per_call tiles (Raster/Job.hs:132-179), in the order
 * buildRasterJobs traverses them (jobs are listed in traversal order; the
 * reference's list is the reverse, CallKernels.hs:255). */
typedef struct gudni_jobs_view {
    int32_t         n_jobs;
    const int32_t  *job_tile_offset;   /* [n_jobs+1]                          */
    const int32_t  *job_shape_offset;  /* [n_jobs+1]                          */
    const uint8_t  *tiles;             /* 32-byte TileInfo records            */
    const uint8_t  *shape_refs;        /* 16-byte Shape records               */
    int32_t         n_tiles, n_refs;
} gudni_jobs_view;

typedef struct gudni_jobs gudni_jobs;
/* Replaces buildTileTree + addShapeToTree + buildRasterJobs
 * (TileTree.hs:81-190, CallKernels.hs:244-255). */
int  gudni_jobs_build(const gudni_spec *spec, const gudni_geometry *g,
                      int32_t canvas_w, int32_t canvas_h, gudni_jobs **out);
int  gudni_jobs_view_get(const gudni_jobs *j, gudni_jobs_view *v);
void gudni_jobs_free(gudni_jobs *j);

/* ---- Level 2: whole frame with every buffer resident in HBM --------------- */
/* Replaces queueRasterJobs (CallKernels.hs:218-242) for a whole frame: the
 * jobs' tiles and shape references are uploaded once, then every frame render
 * runs all tiles of all jobs from device memory in one persistent launch,
 * sorting with the sort_radix_bits of the spec the jobs were built with. */
int gudni_frame_upload(gudni_ctx *ctx, const gudni_jobs *jobs,
                       int32_t width, int32_t height,
                       const float background[4], int32_t frame);
/* Renders the uploaded frame into device memory (stream-ordered, returns
 * without waiting).  out_dev: width*height BGRA words; out_f32_dev (may be
 * NULL): width*height*4 floats, the pre-quantisation RGBA of each pixel. */
int gudni_frame_render(gudni_ctx *ctx, uint32_t *out_dev, float *out_f32_dev);
/* gudni_frame_render without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops): no
 * counterpart in the reference.  The kernels take their threshold rings,
 * shape scratch, control words and scene descriptor from one static device
 * array of the library (at most 34 MiB, held by every GPU the library runs
 * on until the process ends) instead of the context's buffers; the frame
 * must have been uploaded (gudni_frame_upload / gudni_scene_bin, which do
 * allocate).  Limits: at most 131072 work units and max_thresholds at most
 * 4096 (GUDNI_ERR_ARG beyond them); the fast pass runs 32 blocks, so it is
 * slower than gudni_frame_render on a large frame.  The results are those
 * of gudni_frame_render.  Every call on one GPU uses that GPU's array, whose
 * control words are cleared by each call: calls that overlap on one GPU
 * (from any context or stream) race.  Timing is not recorded.  Errors of the
 * kernels surface at gudni_frame_check. */
int gudni_frame_render_static(gudni_ctx *ctx, uint32_t *out_dev, float *out_f32_dev);
/* Waits for the stream and reports a queue overflow or guard hit seen by the
 * last render. */
int gudni_frame_check(gudni_ctx *ctx);

/* ---- Level 2: a whole scene, binned into tiles on the device --------------
 * Replaces buildRasterJobs + queueRasterJobs (OpenCL/CallKernels.hs:218-255)
 * and the tile tree they build (Raster/TileTree.hs:81-204, Raster/Job.hs:
 * 132-179).  The scene holds what the serialiser produces (Raster/
 * Serialize.hs:148-177): the strand heap, one 16-byte Shape record per
 * ShapeEntry in traversal order, and each entry's box (left, top, right,
 * bottom; Raster/Enclosure.hs:62-73).  The device build yields the same
 * tiles, shape references (newest first), job numbering and column
 * allocations as the host build (gudni_jobs_build + gudni_frame_upload). */
typedef struct gudni_scene {
    const void *geo;            size_t geo_bytes;
    const void *shapes;         int32_t n_shapes;      /* 16 B Shape records   */
    const float *bbox;                                 /* 4 floats per shape   */
    const void *substances;     int32_t n_substances;
    const void *pict_data;      size_t pict_bytes;
    const void *pict_uses;      int32_t n_pict_uses;
    const float *random_field;  int32_t n_random;
    int32_t width, height;
    float background[4];
    const void *gradients;      int32_t n_gradients;  /* gudni_gradient records */
} gudni_scene;
/* Uploads and validates the scene (the common buffers of gudni_upload_common
 * plus the shape records and boxes). */
int   gudni_scene_upload(gudni_ctx *ctx, const gudni_scene *scene);
/* Stage (a) on the device: tile tree, per-tile shape references, job fields
 * and the work list, then the frame is ready for gudni_frame_render. */
int   gudni_scene_bin(gudni_ctx *ctx);
/* counts[0] tiles, counts[1] shape references, counts[2] work units. */
int   gudni_scene_counts(gudni_ctx *ctx, int32_t counts[3]);
/* Reads back the device build (any pointer may be NULL): tiles (32 B each,
 * shapeStart absolute into the references), tile_job (reversed job index per
 * tile), references (16 B each), their strand extents (4 floats) and the work
 * list (tile << 2 | quarter). */
int   gudni_scene_read(gudni_ctx *ctx, void *tiles, int32_t *tile_job, void *refs,
                       float *ref_bounds, uint32_t *work);
/* Milliseconds of the last gudni_scene_bin on the device (timing enabled). */
float gudni_last_bin_ms(const gudni_ctx *ctx);

/* Level 2 from outlines: the serialiser runs on the device too.  Replaces
 * the onShape serialisation of the geometry pile (Raster/Serialize.hs:76-177:
 * splitShape -> strands, Raster/Strand.hs:77-206, Deknob.hs:61-108,
 * ReorderTable.hs:44-111; enclose/excludeBox; makeShapeEntry) in addition to
 * what gudni_scene_upload takes over.  `outlines` is the onShape sequence;
 * `scene` supplies the substances, pictures, gradients, random field, canvas
 * and background, and its geo / shapes / bbox fields must be NULL / 0.  The
 * strand heap, Shape records and boxes built on the device are byte-for-byte
 * those gudni_geometry_build builds on the host.  Then gudni_scene_bin and
 * gudni_frame_render as after gudni_scene_upload.  (The heap stays on the
 * device; host-built jobs cannot be validated against it.) */
int   gudni_scene_upload_outlines(gudni_ctx *ctx, const gudni_outlines *outlines,
                                  const gudni_scene *scene);
/* sizes[0] kept shapes, sizes[1] strand heap bytes of the uploaded scene. */
int   gudni_scene_geometry_sizes(gudni_ctx *ctx, int64_t sizes[2]);
/* Reads back the scene's strand heap, Shape records (16 B), boxes and strand
 * extents (4 floats each); any pointer may be NULL. */
int   gudni_scene_geometry_read(gudni_ctx *ctx, void *geo, void *shapes, float *bbox, float *bounds);
/* Milliseconds of the last device serialisation (timing enabled). */
float gudni_last_serialize_ms(const gudni_ctx *ctx);
/* Upload + bin + render + read back, synchronously: the scene's W*H BGRA
 * words into host_out. */
int   gudni_raster_scene(gudni_ctx *ctx, const gudni_scene *scene, uint32_t *host_out);

/* ---- Adjoint: gradients of an image-space L2 loss --------------------------
 * No counterpart in the reference (an extension on the same raster path).
 * Forward: the raster path with GUDNI_COVERAGE_ACCURATE section mid-points and
 * a premultiplied front-to-back composite per section; the image is the
 * pre-quantisation RGBA.  Loss: L = sum over pixels and RGBA channels of
 * (image - target)^2.  Gradients: dL/d(outline points) in the layout of
 * gudni_outlines.pairs (on.x on.y off.x off.y per pair) and dL/d(substance
 * RGBA) for solid substances (zero for pictures and gradients).  Discrete decisions (knob
 * splits, strand runs, tiles, curve subdivision, record drops, section
 * structure) are held; the knob-split parameter follows the analytic vertical
 * tangent.  Shapes may share substances and be subtract compounds: the
 * colour walk merges them as determineColor does (Kernels.cl:1447-1513) and
 * a substance's gradient sums over every layer it colours.  Nonzero-fill
 * shapes (GUDNI_TAG_NONZERO) are refused.  One persistent launch runs the
 * forward and the backward, for all scenes of the handle; the backward
 * sweeps the records the forward left sorted (a unit that overflows its
 * rings regenerates them in the retry pass). */
typedef struct gudni_adjoint_scene {
    gudni_outlines outlines;            /* the geometry parameters           */
    const float *rgba;  int32_t n_substances;   /* 16 B substance records    */
    const void *pict_data;  size_t pict_bytes;
    const void *pict_uses;  int32_t n_pict_uses;
    const float *target;                /* width*height RGBA floats          */
    int32_t width, height;
    float background[4];
    const void *gradients;  int32_t n_gradients;   /* gudni_gradient records */
} gudni_adjoint_scene;
typedef struct gudni_adjoint_batch gudni_adjoint_batch;
/* Serialises, bins and uploads n_scenes scenes with their targets. */
int  gudni_adjoint_upload(gudni_ctx *ctx, int32_t n_scenes, const gudni_adjoint_scene *scenes,
                          gudni_adjoint_batch **out);
/* New parameters for one scene (pairs and/or rgba may be NULL = unchanged),
 * from host memory: copied up, then the scene is re-serialised (with the
 * curve provenance the point gradients need) and re-binned on the device,
 * and the batch's tables are reassembled; synchronous.  The first update of
 * a batch rebuilds every scene this way; the result is identical to the
 * host build gudni_adjoint_upload makes.  If an update fails the batch
 * refuses to run until uploaded again. */
int  gudni_adjoint_update(gudni_ctx *ctx, gudni_adjoint_batch *a, int32_t scene, const float *pairs,
                          const float *rgba);
/* The same with the new parameters in device memory (e.g. a training loop's
 * tensors): no host round trip of the parameters.  For both update calls,
 * scene -1 updates every scene at once from the concatenation of all
 * scenes' pairs (and RGBA) in batch order, reassembling the tables once. */
int  gudni_adjoint_update_device(gudni_ctx *ctx, gudni_adjoint_batch *a, int32_t scene,
                                 const float *pairs, const float *rgba);
/* Forward + backward of every scene in one launch (stream-ordered). */
int  gudni_adjoint_run(gudni_ctx *ctx, gudni_adjoint_batch *a);
/* The same for one scene alone. */
int  gudni_adjoint_run_scene(gudni_ctx *ctx, gudni_adjoint_batch *a, int32_t scene);
/* Waits and copies one scene's loss, gradients and image (any may be NULL). */
int  gudni_adjoint_result(gudni_ctx *ctx, gudni_adjoint_batch *a, int32_t scene, double *loss,
                          float *d_pairs, float *d_rgba, float *image);
/* The same into device memory, stream-ordered, without waiting (scene -1:
 * every scene's losses, gradients and images, concatenated); kernel
 * errors (queue overflow, section guard) surface at the next
 * gudni_adjoint_result or gudni_adjoint_check. */
int  gudni_adjoint_result_device(gudni_ctx *ctx, gudni_adjoint_batch *a, int32_t scene, double *loss,
                                 float *d_pairs, float *d_rgba, float *image);
/* Waits for the stream and reports the last run's kernel errors. */
int  gudni_adjoint_check(gudni_ctx *ctx, gudni_adjoint_batch *a);
void gudni_adjoint_free(gudni_adjoint_batch *a);
/* One scene, synchronously: upload, run, result, free. */
int  gudni_adjoint(gudni_ctx *ctx, const gudni_adjoint_scene *scene, double *loss, float *d_pairs,
                   float *d_rgba, float *image);
/* Pass timings (ms) of the last render, recorded with hipEvents on the
 * launch stream when enabled and read by gudni_frame_check: [0] counter
 * reset + fast pass, [1] retry pass. */
/* Launch shape of the persistent fast pass: blocks (of four waves) per
 * compute unit, 1..8, or 0 for as many as fit (at most 8).  Fewer resident
 * waves keep more of each wave's threshold rings in L2; more hide more
 * latency.  No counterpart in the reference (its launch shape is fixed by
 * the OpenCL runtime). */
int gudni_set_blocks_per_cu(gudni_ctx *ctx, int32_t blocks_per_cu);
/* Byte order of the 32-bit pixel words every render of the context writes:
 * GUDNI_FORMAT_BGRA8 (default) is the reference's word, bytes B, G, R, A in
 * memory (colorToSolidPixel_Word32_BGRA, Kernels.cl:842-844; the ARGB8888
 * texture of Interface/InterfaceSDL.hs:118); GUDNI_FORMAT_RGBA8 stores the
 * same channel values as bytes R, G, B, A (an RGBA8 image).  Alpha is 255
 * in both.  A batch's captured graph is re-captured when this changes. */
#define GUDNI_FORMAT_BGRA8 0
#define GUDNI_FORMAT_RGBA8 1
int gudni_set_output_format(gudni_ctx *ctx, int32_t format);
int gudni_enable_timing(gudni_ctx *ctx, int enable);
int gudni_last_timings(gudni_ctx *ctx, float *ms, int32_t n);
/* Shader-clock cycles of the last timed render's fast pass, summed over
 * waves: [0] threshold generation, [1] sort, [2] coverage/compositing sweep,
 * [3] work units processed.  After a timed gudni_adjoint_run and its
 * gudni_adjoint_result (up to 8 entries): the forward's generation, sort and
 * sweep, the backward's generation, sort and sweep, the pass that moves line
 * weights onto outline points, and work units. */
int gudni_last_stage_cycles(gudni_ctx *ctx, uint64_t *cycles, int32_t n);

/* ---- Batches: many independent scenes in one persistent launch ----------- */
/* No counterpart in the reference (it renders one scene per frame, one job
 * per launch, CallKernels.hs:218-242).  Each scene brings the buffers
 * gudni_upload_common takes plus its jobs; the batch concatenates them in
 * device memory with one frame per scene, and gudni_batch_render runs every
 * tile of every scene through the same kernels in one persistent launch
 * (optionally replayed from a captured hipGraph). */
typedef struct gudni_batch_scene {
    const void *geo;          size_t geo_bytes;
    const void *substances;   int32_t n_substances;
    const void *pict_data;    size_t pict_bytes;
    const void *pict_uses;    int32_t n_pict_uses;
    const gudni_jobs *jobs;
    int32_t width, height;
    float background[4];
    const void *gradients;    int32_t n_gradients;    /* gudni_gradient records */
} gudni_batch_scene;

typedef struct gudni_batch gudni_batch;
int  gudni_batch_upload(gudni_ctx *ctx, int32_t n_scenes, const gudni_batch_scene *scenes,
                        const float *random_field, int32_t n_random, gudni_batch **out);
/* Each scene's records are sorted with the sort_radix_bits of the spec its
 * jobs were built with (so a tuner may choose tile size and radix per
 * scene). */
/* The same batch built from outlines on the device: each scene is
 * serialised (as gudni_scene_upload_outlines) and binned (as
 * gudni_scene_bin) with its own tile size tile_sizes[s] (a power of two in
 * 8..256; NULL: the spec's max_tile_size) and sort radix radix_bits[s]
 * (0..5; NULL: the spec's), then appended to the batch's device arrays.
 * scenes[s] supplies canvas, background, substances, pictures and
 * gradients; its geo and jobs must be empty.  Reuses the context's
 * uploaded-frame and scene buffers, so a frame uploaded before must be
 * uploaded again.  No counterpart in the reference, whose serialiser and
 * tile builder run on the host for one frame at a time (Raster/Serialize.hs:
 * 148-177, Raster/TileTree.hs:81-204, Raster/Job.hs:132-179). */
int  gudni_batch_upload_outlines(gudni_ctx *ctx, int32_t n_scenes, const gudni_outlines *outlines,
                                 const gudni_batch_scene *scenes, const int32_t *tile_sizes,
                                 const int32_t *radix_bits, const float *random_field,
                                 int32_t n_random, gudni_batch **out);
/* Renders every scene (stream-ordered).  use_graph: on the first call the
 * launch sequence is captured into a hipGraph on a stream the batch owns,
 * then that graph is replayed on the context's stream (the null stream
 * included) by this and every later call. */
int  gudni_batch_render(gudni_ctx *ctx, gudni_batch *b, int use_graph);
/* Renders one scene of the batch by itself (the sequential-launch baseline). */
int  gudni_batch_render_scene(gudni_ctx *ctx, gudni_batch *b, int32_t scene);
/* Waits, reports overflow/guard errors, and fills latency_ms[n_scenes]
 * (may be NULL): device time from the start of the last gudni_batch_render
 * to the end of each scene's last work unit. */
int  gudni_batch_check(gudni_ctx *ctx, gudni_batch *b, float *latency_ms);
/* Copies one scene's BGRA frame (width*height words) to host memory. */
int  gudni_batch_frame(gudni_ctx *ctx, gudni_batch *b, int32_t scene, uint32_t *host_out);
void gudni_batch_free(gudni_batch *b);
/* The order a whole-batch launch hands out its work units (tile << 2 |
 * quarter, over `tiles`, TileInfo records).  A unit's estimated cost is the
 * rows each of its threads sweeps x ((the tile's shape count + 1) / tile
 * width)^2; the units within a factor 8 of the costliest go first, longest
 * first, dealt tile by tile, round robin, to the starts of the `parts`
 * contiguous ranges the launch's claimers begin with (range p = [n p / parts,
 * n (p+1) / parts)), so that none is claimed last.  With `tile_scene` (the
 * scene of each tile; may be null) the others follow scene by scene, the
 * scene with the smallest summed cost first, each scene's units together and
 * in their order, scenes dealt round robin to the ranges (a full range passes
 * the rest on to the next), so small scenes finish early; without it they
 * follow in their original order.  Batches built by gudni_batch_upload and
 * gudni_batch_upload_outlines use it with their scenes; per-scene launches
 * keep each scene's own order.  An extension: the reference launches one job
 * at a time (OpenCL/CallKernels.hs:244-255). */
int  gudni_order_work(const void *tiles, int32_t n_tiles, const int32_t *tile_scene, const uint32_t *work,
                      int32_t n_work, int32_t threads_per_tile, int32_t parts, uint32_t *out);

/* Test hook: per-column threshold records after generate (queue order) and
 * after sort, plus the initial shape state, for one RasterJob.  Slot =
 * tile_in_job * threads_per_tile + column; `cap` records are kept per slot,
 * front first.  All arrays are host memory sized by the caller. */
typedef struct gudni_dump {
    int32_t   cap;
    float    *pre_thresholds;   /* [slots][cap][4]                            */
    uint32_t *pre_headers;      /* [slots][cap]                               */
    float    *thresholds;       /* [slots][cap][4]                            */
    uint32_t *headers;          /* [slots][cap]                               */
    int32_t  *queue_len;        /* [slots] (-1: inactive column)              */
    uint32_t *shape_bits;       /* [slots]                                    */
    uint64_t *shape_stack;      /* [slots][2]                                 */
    int32_t  *shape_indices;    /* [slots][max_shapes]                        */
} gudni_dump;
int gudni_dump_job(gudni_ctx *ctx,
                   const void *shapes, int32_t n_shapes,
                   const void *tiles, int32_t n_tiles,
                   int32_t width, int32_t height, int32_t job_index,
                   gudni_dump *dump);

#ifdef __cplusplus
}
#endif
#endif
