/*
 * gudni_oracle.h -- CPU restatement of Gudni's raster path, used ONLY as the
 * correctness checker (tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg).  Nothing in the gudni_amd product links or loads it.
 *
 * It restates, in plain C with IEEE binary32 arithmetic and no contraction:
 *   - the strand serialiser   (Raster/Strand.hs, Deknob.hs, ReorderTable.hs,
 *                              Enclosure.hs, Serialize.hs:76-177)
 *   - the tile tree and jobs  (Raster/TileTree.hs:81-204, Raster/Job.hs:132-179,
 *                              OpenCL/CallKernels.hs:244-255)
 *   - the three device kernels (OpenCL/Kernels.cl:1131-2167) with a per-column
 *     threshold queue laid out as in the reference.
 * plus an FP64 shadow of the coverage integration (same sections, areas and
 * accumulation in double).
 */
#ifndef GUDNI_ORACLE_H
#define GUDNI_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct go_spec {
    int32_t max_thresholds;       /* MAXTHRESHOLDS  (Setup.hs:49)              */
    int32_t max_shapes;           /* MAXSHAPE = 127 (Constants.hs:58)          */
    int32_t compute_depth;        /* log2 threadsPerTile (CallKernels.hs:121)  */
    int32_t threads_per_tile;     /* 256 (Setup.hs:82)                         */
    int32_t tiles_per_call;       /* 256 (Setup.hs:83)                         */
    int32_t max_tile_size;        /* 256 (Setup.hs:81)                         */
    int32_t max_strands_per_tile; /* maxThresholds - 2 (Setup.hs:85)           */
    int32_t random_field_size;    /* RANDOMFIELDSIZE 4096 (Constants.hs:67)    */
    float   stochastic_factor;    /* STOCHASTIC_FACTOR 0.0 (Constants.hs:54)   */
    int32_t coverage_mode;        /* 0: section mid-x evaluated relative to the
                                     threshold start (accurate); 1: in the
                                     reference's order (Kernels.cl:916-926)   */
    int32_t texture_filter;       /* 0: nearest texel (the reference's only
                                     mode, Kernels.cl:856-860); 1: bilinear   */
} go_spec;

/* Strand serialisation ---------------------------------------------------- */
int  go_reorder_entry(int size, int i);
/* Serialise the outlines of one shape (Enclosure.hs:62-73).  pairs: 4 floats
 * per curve pair.  Returns bytes written (multiple of 16) or -1 if cap is too
 * small.  *n_strands receives the strand count. */
long go_serialize_outlines(const float *pairs, const int32_t *outline_offsets,
                           int32_t n_outlines, uint8_t *out, long cap,
                           int32_t *n_strands);
/* Knob split of one quadratic (Deknob.hs:61-102).  Returns 1 or 2 curves,
 * written as 6 floats each. */
int  go_fix_knob(const float *bez6, float *out12);

/* Whole-scene serialisation (Serialize.hs:148-177): culled shapes are
 * dropped; for kept shapes writes 16-byte shape records, bbox and strand
 * counts in traversal order. */
typedef struct go_serialized {
    uint8_t  *geo;          long geo_bytes;
    uint8_t  *shapes;       /* 16 B records                                 */
    float    *bbox;         /* [n][4] left top right bottom                 */
    uint32_t *strand_count;
    int32_t  *source_index; /* index in the input scene                     */
    int32_t   n_shapes;
} go_serialized;
int  go_serialize_scene(int32_t n_shapes, const uint64_t *tags,
                        const int32_t *shape_outline_offsets,
                        const int32_t *outline_pair_offsets,
                        const float *pairs, float canvas_w, float canvas_h,
                        go_serialized *out);
void go_free_serialized(go_serialized *s);

/* Tile tree + jobs, literal insert/split procedure (TileTree.hs:113-190). */
typedef struct go_jobs {
    int32_t  n_jobs;
    int32_t *job_tile_offset;   /* [n_jobs+1] into tiles                  */
    int32_t *job_shape_offset;  /* [n_jobs+1] into shape refs             */
    uint8_t *tiles;             /* 32-byte TileInfo records               */
    uint8_t *shape_refs;        /* 16-byte Shape records                  */
    int32_t  n_tiles, n_refs;
} go_jobs;
int  go_build_jobs(const go_spec *spec, int32_t canvas_w, int32_t canvas_h,
                   const go_serialized *scene, go_jobs *out);
void go_free_jobs(go_jobs *j);

/* One raster job (the three kernels) ------------------------------------- */
typedef struct go_job_in {
    const uint8_t *geo;          /* strand heap (float4-addressable)        */
    const uint8_t *shapes;       /* 16-byte Shape records of the job        */
    const uint8_t *tiles;        /* 32-byte TileInfo records of the job     */
    int32_t        n_tiles;
    const float   *substances;   /* 4 floats each                           */
    int32_t        n_substances;
    const uint8_t *pict_data;    /* RGBA8 pile                              */
    long           pict_bytes;
    const uint8_t *pict_uses;    /* 24-byte PictureUse records              */
    int32_t        n_pict_uses;
    const float   *random_field; /* random_field_size floats                */
    float          background[4];
    int32_t        width, height, frame, job_index;
    const float   *gradients;    /* 28 floats (112 B) per gradient record    */
    int32_t        n_gradients;
} go_job_in;

/* Optional dumps for parity tests.  Column slot = tile_in_job*tpt + column.
 * Records are stored front-first, `cap` per slot; queue_len may exceed cap
 * (then only the first cap records are kept). */
typedef struct go_job_dump {
    int32_t   cap;
    float    *pre_thresholds;  /* [slot][cap][4] queue order after generate */
    uint32_t *pre_headers;     /* [slot][cap]                               */
    float    *thresholds;      /* [slot][cap][4] after sort                 */
    uint32_t *headers;         /* [slot][cap]                               */
    int32_t  *queue_len;       /* [slot], -1 for inactive columns           */
    uint32_t *shape_bits;      /* [slot]                                    */
    uint64_t *shape_stack;     /* [slot][2] initial stack (bits 0..127)     */
    int32_t  *shape_indices;   /* [slot][max_shapes]                        */
} go_job_dump;

/* Returns 0, -1 on queue overflow (the reference corrupts the neighbouring
 * column there) or -2 when a column exceeds the section-count guard.  out: W*H BGRA words (only the job's pixels are written);
 * out_f32 / out_f64 optional W*H*4 pre-quantisation colours (RGBA). */
int go_raster_job(const go_spec *spec, const go_job_in *in, uint32_t *out,
                  float *out_f32, double *out_f64, go_job_dump *dump);

/* Convenience: serialise + bin + raster every job (host path of
 * Application.drawFrame).  nthreads > 1 spreads tiles over threads. */
int go_render_scene(const go_spec *spec, const go_serialized *scene,
                    const go_jobs *jobs, const go_job_in *common,
                    uint32_t *out, float *out_f32, double *out_f64,
                    int nthreads);

void go_default_spec(go_spec *spec);

/* Test hook: the ordinate of the column border (x = dx for the left border,
 * dx + 1 for the right) that one strand contributes in buildThresholdArray
 * (searchTree + bifurcateCurve + yIntercept, Kernels.cl:1226-1408), in
 * column-local coordinates.  Returns 0 if the strand misses the column. */
int go_strand_border(const uint8_t *strand, float dx, float dy, int is_left,
                     float *y, int *halvings, int *node);

/* FP64 restatement of the differentiable forward (gudni_diff.c): the loss
 * sum (pixel - target)^2 over the window [x0,x1) x [y0,y1) of a scene of
 * solid substances (rgba: 4 doubles per substance), the window's pixels
 * (optional, row-major RGBA doubles) and a signature of every discrete
 * decision.  Returns 0, or < 0 on queue overflow / sweep guard. */
int go_diff_loss(int32_t n_shapes, const uint64_t *tags, const int32_t *shape_outline_offsets,
                 const int32_t *outline_pair_offsets, const double *pairs, const double *rgba,
                 const double *background, int32_t width, int32_t height, const float *target,
                 int32_t x0, int32_t y0, int32_t x1, int32_t y1, int32_t seg, double *loss,
                 double *image, uint64_t *signature);

/* go_diff_loss with (a) seg_map (may be NULL): W*H int32, the height of the
 * column segment each pixel belongs to (powers of two, segments at
 * multiples of it, as a tile tree of mixed tile sizes cuts the columns):
 * each column of the window runs over the whole segments holding its rows;
 * and (b) picture substances: substances (raw 16-byte records; a picture's
 * first word is its PictureUse index), the picture pile and PictureUse
 * records, sampled nearest (filter 0) or bilinear (1) at the pixel's
 * integer position.  substances NULL: solid colours only.  Returns -4 for
 * gradient substances or nonzero-fill shapes (not modelled). */
int go_diff_loss_ex(int32_t n_shapes, const uint64_t *tags, const int32_t *shape_outline_offsets,
                    const int32_t *outline_pair_offsets, const double *pairs, const double *rgba,
                    const double *background, int32_t width, int32_t height, const float *target,
                    int32_t x0, int32_t y0, int32_t x1, int32_t y1, int32_t seg,
                    const int32_t *seg_map, const float *substances, const uint8_t *pict_data,
                    const uint8_t *pict_uses, int32_t n_pict_uses, int32_t filter,
                    double *loss, double *image, uint64_t *signature);

/* The adjoint's algorithm restated in FP64 over the whole canvas, columns
 * cut into segments of `seg` rows: loss, dL/d(outline points) (4 doubles per
 * pair) and dL/d(substance RGBA). */
int go_diff_grad(int32_t n_shapes, const uint64_t *tags, const int32_t *shape_outline_offsets,
                 const int32_t *outline_pair_offsets, const double *pairs, const double *rgba,
                 const double *background, int32_t width, int32_t height, const float *target,
                 int32_t seg, double *loss, double *d_pairs, double *d_rgba, int32_t n_pairs,
                 int32_t n_substances);

#ifdef __cplusplus
}
#endif
#endif
