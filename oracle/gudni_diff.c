/*
 * gudni_diff.c -- FP64 restatement of the differentiable forward, for
 * finite-difference checks of the adjoint.
 *
 * TEST INFRASTRUCTURE ONLY (see gudni_oracle.c): only tests/ load it.
 *
 * It recomputes, in double precision and straight from the outline points,
 * what the raster path computes for a window of pixel columns:
 *   outline -> quadratics, knobs split at the exact vertical tangent
 *     t* = (p0.x - c.x) / (p0.x - 2c.x + p1.x)      (Raster/Deknob.hs:61-108)
 *   -> x-monotone strands of <= 16 curves, the final run first, runs that go
 *     right-to-left reversed                        (Raster/Strand.hs:77-178)
 *   -> per column: the curves found at the column's left and right edges
 *     (searchTree's left- and right-biased descent, Kernels.cl:1337-1357,
 *     restated as a count of interior points), wings and bridge
 *     (spawnThresholds :1264-1333), border intercepts by the same midpoint
 *     bisection to taxicab flatness 0.25 (:1226-1258), records with the
 *     MINCROP drop and top trim (:1131-1220), pushed at the queue front
 *   -> stable sort by (top, topX, inverse slope)     (:1079-1094, 1932-1976)
 *   -> section sweep (verticalAdvance / splitNext / insertThreshold /
 *     horizontalAdvance, :1007-1124, 1744-1851) with the section mid-x taken
 *     at mid-height of the span, and per section the premultiplied
 *     front-to-back composite of the set shapes (front = lowest scene index,
 *     stopping at alpha 1) over the background, straightened
 *   -> pixel = sum(area * colour) / sum(area), loss = sum (pixel - target)^2.
 * A column segment here is the window's rows (go_diff_loss) or `seg` rows
 * aligned to multiples of seg (go_diff_grad).  Tiles cut columns into such
 * segments; each segment trims the records crossing its top and sorts its
 * own, and since the sweep never reorders active records that cross inside a
 * span (a section of negative width follows), the forward depends on where
 * the columns are cut: checks of a tiled forward sum go_diff_loss over the
 * same segments.
 *
 * Every discrete decision (culling, knob splits, strand runs, the curves
 * found, bisection paths, record drops, the number of sections per pixel)
 * is folded into a 64-bit signature so a finite difference can be rejected
 * when a decision flips between its two evaluations.
 */
#include "gudni_oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define D_MINCROP 0.2
#define D_FLAT 0.25
#define D_MAX_BISECT 64
#define D_QCAP (1 << 16)
#define D_TAG_ID 0x07FFFFFFFFFFFFFFull
#define D_TAG_TYPE 0xC000000000000000ull
#define D_TAG_COMPOUND 0x3000000000000000ull
#define D_TAG_ADD 0x2000000000000000ull
#define D_TAG_CONTINUE 0x1000000000000000ull
#define D_TAG_PICTURE 0x4000000000000000ull
#define D_POS 0x80000000u
#define D_PERSIST 0x40000000u
#define D_PERSIST_AND_SLOPE 0xC0000000u
#define D_PERSIST_TOP 0xC0000000u
#define D_PERSIST_BOTTOM 0x40000000u

typedef struct { double x, y; } d2;
typedef struct { d2 a, c, b; } dq;                 /* start, control, end */
typedef struct { double top, bottom, left, right; uint32_t h; int label; int origin; } drec;

static uint64_t fnv(uint64_t s, uint64_t v)
{
    for (int i = 0; i < 8; i++) { s ^= (v >> (8 * i)) & 0xFF; s *= 0x100000001b3ull; }
    return s;
}

/* ---- serialisation ------------------------------------------------------ */
/* where a strand curve comes from: the outline curve's point indices (pair
 * q's on-curve point is 2q, its control 2q+1), the knob-split parameter, the
 * half (1 lower, 2 upper) and whether the strand runs it backwards (4) */
typedef struct { int p0, c, p1; double t; int flags; } dsrc;

typedef struct {
    int kept;
    int n_strands;
    int *strand_start;    /* [n_strands + 1] into curves */
    dq *curves;
    dsrc *src;            /* per curve */
    uint64_t sid;
    uint64_t tag;
    double xmin, xmax;    /* strands' left-most and right-most end points */
    double ymin, ymax;    /* over every curve point (end points and controls) */
} dshape;

static d2 lerp2(double t, d2 p, d2 q) { d2 r = { p.x * (1 - t) + q.x * t, p.y * (1 - t) + q.y * t }; return r; }
static int xdir(const dq *q) { return q->a.x < q->b.x ? -1 : (q->a.x > q->b.x ? 1 : 0); }

static void build_shape(dshape *s, const double *pairs, const int32_t *ooff, int o0, int o1,
                        int W, int H, uint64_t *sig)
{
    s->kept = 0; s->n_strands = 0; s->strand_start = NULL; s->curves = NULL; s->src = NULL;
    int p0 = ooff[o0], p1 = ooff[o1];
    if (p1 <= p0) return;
    double l = pairs[4 * p0], t = pairs[4 * p0 + 1], r = l, b = t;
    for (int p = p0; p < p1; p++)
        for (int h = 0; h < 2; h++) {
            double x = pairs[4 * p + 2 * h], y = pairs[4 * p + 2 * h + 1];
            if (x < l) l = x;
            if (x > r) r = x;
            if (y < t) t = y;
            if (y > b) b = y;
        }
    if (l >= W || t >= H || r <= 0.0 || b <= 0.0) { *sig = fnv(*sig, 0xC0111ull); return; }
    s->kept = 1;
    int cap = 2 * (p1 - p0) + 16;
    dq *work = malloc(sizeof(dq) * cap);
    dsrc *wsrc = malloc(sizeof(dsrc) * cap);
    s->curves = malloc(sizeof(dq) * cap);
    s->src = malloc(sizeof(dsrc) * cap);
    s->strand_start = malloc(sizeof(int) * (cap + 1));
    int nc = 0;
    for (int o = o0; o < o1; o++) {
        int a = ooff[o], e = ooff[o + 1], k = e - a, nw = 0;
        if (k < 2) continue;
        for (int i = 0; i < k; i++) {
            int j = (i + 1 == k) ? 0 : i + 1;
            dq q = { { pairs[4 * (a + i)], pairs[4 * (a + i) + 1] },
                     { pairs[4 * (a + i) + 2], pairs[4 * (a + i) + 3] },
                     { pairs[4 * (a + j)], pairs[4 * (a + j) + 1] } };
            int bl = q.c.x < q.a.x && q.c.x < q.b.x, br = q.c.x > q.a.x && q.c.x > q.b.x;
            *sig = fnv(*sig, (uint64_t)(bl | (br << 1)));
            dsrc sr = { 2 * (a + i), 2 * (a + i) + 1, 2 * (a + j), 0.0, 0 };
            if (bl || br) {
                double ts = (q.a.x - q.c.x) / (q.a.x - 2.0 * q.c.x + q.b.x);
                d2 m0 = lerp2(ts, q.a, q.c), m1 = lerp2(ts, q.c, q.b), on = lerp2(ts, m0, m1);
                dq u = { q.a, m0, on }, v = { on, m1, q.b };
                sr.t = ts;
                sr.flags = 1; wsrc[nw] = sr; work[nw++] = u;
                sr.flags = 2; wsrc[nw] = sr; work[nw++] = v;
            } else { wsrc[nw] = sr; work[nw++] = q; }
        }
        int *cuts = malloc(sizeof(int) * (nw + 2)), nr = 0;
        cuts[nr++] = 0;
        for (int i = 1; i < nw; i++) {
            int da = xdir(&work[i - 1]), db = xdir(&work[i]);
            if (da != db || da == 0) cuts[nr++] = i;
        }
        cuts[nr] = nw;
        int runs = nr;
        for (int rr = 0; rr < runs; rr++) {
            int run = (rr == 0) ? runs - 1 : rr - 1;
            for (int st = cuts[run]; st < cuts[run + 1]; st += 16) {
                int m = cuts[run + 1] - st; if (m > 16) m = 16;
                const dq *src = work + st;
                int rev = src[0].a.x > src[m - 1].b.x;
                s->strand_start[s->n_strands++] = nc;
                for (int q = 0; q < m; q++) {
                    if (rev) {
                        const dq *w = &src[m - 1 - q]; dq z = { w->b, w->c, w->a };
                        s->src[nc] = wsrc[st + m - 1 - q]; s->src[nc].flags |= 4;
                        s->curves[nc++] = z;
                    } else { s->src[nc] = wsrc[st + q]; s->curves[nc++] = src[q]; }
                }
                *sig = fnv(*sig, (uint64_t)(m * 2 + rev));
            }
        }
        free(cuts);
    }
    s->strand_start[s->n_strands] = nc;
    s->xmin = DBL_MAX; s->xmax = -DBL_MAX;
    s->ymin = DBL_MAX; s->ymax = -DBL_MAX;
    for (int i = 0; i < nc; i++) {
        const double ys[3] = { s->curves[i].a.y, s->curves[i].c.y, s->curves[i].b.y };
        for (int k = 0; k < 3; k++) {
            if (ys[k] < s->ymin) s->ymin = ys[k];
            if (ys[k] > s->ymax) s->ymax = ys[k];
        }
    }
    for (int k = 0; k < s->n_strands; k++) {
        const double lx = s->curves[s->strand_start[k]].a.x, rx = s->curves[s->strand_start[k + 1] - 1].b.x;
        if (lx < s->xmin) s->xmin = lx;
        if (rx > s->xmax) s->xmax = rx;
    }
    free(work);
    free(wsrc);
}

/* ---- generation --------------------------------------------------------- */
typedef struct {
    drec *q; int qstart, qlen, overflow;
    int height_i; double fh;
    uint64_t *stack; int nwords;
} dcol;

typedef struct { d2 right, left, ctl; double xpos; int index; } dtrav;

static void push_front(dcol *c, drec r)
{
    if (c->qstart == 0) { c->overflow = 1; return; }
    c->qstart--; c->qlen++;
    c->q[c->qstart] = r;
}

static double inv_slope(const drec *t)
{
    double sgn = (t->h & D_POS) ? 1.0 : -1.0;
    return (t->top == t->bottom) ? DBL_MAX : (t->right - t->left) / (t->bottom - t->top) * sgn;
}
static double top_x(const drec *t) { return (t->h & D_POS) ? t->left : t->right; }
static double isect_x(const drec *t, double y)
{
    if (t->left == t->right) return t->left;
    double sy = (t->h & D_POS) ? t->top : t->bottom;
    return (y - sy) * inv_slope(t) + t->left;
}
static int is_below(const drec *a, const drec *b)
{
    if (a->top > b->top) return 1;
    if (a->top == b->top) {
        double ax = top_x(a), bx = top_x(b);
        if (ax > bx) return 1;
        if (ax == bx && inv_slope(a) > inv_slope(b)) return 1;
    }
    return 0;
}

static void add_line(dcol *c, d2 l, d2 r, int label, int *added, int *enclosed, uint64_t *sig)
{
    drec t;
    t.top = l.y < r.y ? l.y : r.y;    /* min(l.y, r.y) */
    t.bottom = l.y < r.y ? r.y : l.y;
    t.left = l.x; t.right = r.x; t.label = label;
    int pos = l.y <= r.y;
    int persistent = (l.x != r.x) && (l.x == 0.0);
    t.h = (pos ? D_POS : 0u) | (persistent ? D_PERSIST : 0u);
    int keep = persistent || (t.bottom - t.top) >= D_MINCROP;
    *sig = fnv(*sig, (uint64_t)keep);
    if (!keep) return;
    *enclosed = *enclosed || ((t.top <= 0.0) && (t.h & D_PERSIST_AND_SLOPE) == D_PERSIST_TOP)
                          || ((t.bottom <= 0.0) && (t.h & D_PERSIST_AND_SLOPE) == D_PERSIST_BOTTOM);
    if (t.top < c->fh && t.bottom > 0.0 && t.left < 1.0) {
        if (t.top <= 0.0) {
            double sx = isect_x(&t, 0.0);
            if (pos) { t.top = 0.0; t.left = sx; t.h &= ~D_PERSIST; }
            else { t.top = 0.0; t.right = sx; }
        }
        if (t.right <= 0.0) { *enclosed = 1; *sig = fnv(*sig, 7); }
        else { *added = 1; push_front(c, t); *sig = fnv(*sig, 9); }
    }
}

static double intersect(dtrav t, uint64_t *sig)
{
    int steps = 0;
    if (!(t.left.x == t.ctl.x && t.left.y == t.ctl.y)) {
        for (int it = 0; it < D_MAX_BISECT; it++) {
            d2 lm = { 0.5 * t.left.x + 0.5 * t.ctl.x, 0.5 * t.left.y + 0.5 * t.ctl.y };
            d2 rm = { 0.5 * t.ctl.x + 0.5 * t.right.x, 0.5 * t.ctl.y + 0.5 * t.right.y };
            d2 on = { 0.5 * lm.x + 0.5 * rm.x, 0.5 * lm.y + 0.5 * rm.y };
            double flat = fabs(t.ctl.x - on.x) + fabs(t.ctl.y - on.y);
            if (!(flat > D_FLAT)) break;
            int go_left = t.xpos < on.x;
            *sig = fnv(*sig, (uint64_t)(go_left + 2));
            if (go_left) { t.ctl = lm; t.right = on; } else { t.left = on; t.ctl = rm; }
            steps++;
        }
    }
    *sig = fnv(*sig, (uint64_t)steps);
    if (t.left.x == t.right.x) return t.left.y < t.right.y ? t.left.y : t.right.y;
    return ((t.right.y - t.left.y) / (t.right.x - t.left.x)) * (t.xpos - t.left.x) + t.left.y;
}

static void spawn(dcol *c, int label, const dtrav *l, const dtrav *r, int *added, int *enclosed, uint64_t *sig)
{
    double yl = (l->left.x >= 0.0) ? l->left.y : intersect(*l, sig);
    int lw = 0, rw = 0;
    if (l->right.x < 1.0 && l->right.x > 0.0) {
        d2 a = { l->xpos, yl };
        add_line(c, a, l->right, label, added, enclosed, sig); lw = 1;
    }
    double yr = (r->right.x <= 1.0) ? r->right.y : intersect(*r, sig);
    if (r->left.x > 0.0 && r->left.x < 1.0 && l->index != r->index) {
        d2 b = { r->xpos, yr };
        add_line(c, r->left, b, label, added, enclosed, sig); rw = 1;
    }
    if (l->right.x < r->left.x || (!lw && !rw)) {
        d2 bl, br;
        if (lw || l->left.x == l->right.x) bl = l->right; else { bl.x = l->xpos; bl.y = yl; }
        if (rw || r->left.x == r->right.x) br = r->left; else { br.x = r->xpos; br.y = yr; }
        add_line(c, bl, br, label, added, enclosed, sig);
    }
}

/* the curve of a strand (curves cs[0..m-1], left to right) holding x:
 * left-biased counts interior points strictly left of x, right-biased those
 * at or left of x */
static void find_curve(dtrav *t, const dq *cs, int m, double dx, double dy, int is_left)
{
    int i = 0;
    for (int k = 1; k < m; k++) {
        double px = cs[k].a.x - dx;
        if (is_left ? (px < t->xpos) : (px <= t->xpos)) i = k;
        else break;
    }
    t->index = i;
    t->left.x = cs[i].a.x - dx; t->left.y = cs[i].a.y - dy;
    t->ctl.x = cs[i].c.x - dx; t->ctl.y = cs[i].c.y - dy;
    t->right.x = cs[i].b.x - dx; t->right.y = cs[i].b.y - dy;
}

static void flip(dcol *c, int label) { c->stack[label >> 6] ^= 1ull << (label & 63); }

/* ---- sweep -------------------------------------------------------------- */
typedef struct { int current, num_active; double sx0, sy0, sx1, sy1, pixel_y; } dsweep;

static void insert_rec(dcol *c, drec t)
{
    if (c->qstart == 0) { c->overflow = 1; return; }
    c->qstart--; c->qlen++;
    int cur = 0;
    while (cur < c->qlen - 1) {
        drec *o = &c->q[c->qstart + cur + 1];
        if (!is_below(&t, o)) break;
        c->q[c->qstart + cur] = *o;
        cur++;
    }
    c->q[c->qstart + cur] = t;
}

/* go_diff_loss folds split decisions into its signature; one per thread, so
 * losses may be evaluated in parallel threads */
static _Thread_local uint64_t *g_split_sig;

static double split_next(dcol *c, dsweep *p)
{
    double slice = DBL_MAX, top = c->q[c->qstart].top;
    int n = 1;
    while (n < c->qlen) { double nt = c->q[c->qstart + n].top; if (nt > top) { slice = nt; break; } n++; }
    p->num_active = n;
    double sp = slice;
    for (int i = 0; i < n; i++) { double b = c->q[c->qstart + i].bottom; if (top < b && b < sp) sp = b; }
    if (sp < slice) slice = sp;
    for (int cur = 0; cur < n && !c->overflow; cur++) {
        drec t = c->q[c->qstart + cur];
        if (t.top < slice && slice < t.bottom) {
            double sx = isect_x(&t, slice);
            drec b = t, a = t;
            if (t.h & D_POS) {
                b.top = slice; b.left = sx; b.h &= ~D_PERSIST;
                a.bottom = slice; a.right = sx;
            } else {
                b.top = slice; b.right = sx;
                a.bottom = slice; a.left = sx; a.h &= ~D_PERSIST;
            }
            c->q[c->qstart + cur] = a;
            const int keep = (b.h & D_PERSIST) || (b.bottom - b.top) >= D_MINCROP;
            if (g_split_sig) *g_split_sig = fnv(*g_split_sig, (uint64_t)(keep + 16));
            if (keep) insert_rec(c, b);
        }
    }
    return slice;
}

static void vertical_advance(dcol *c, dsweep *p)
{
    if (p->sx1 != 1.0) return;
    for (int i = 0; i < p->num_active; i++) flip(c, c->q[c->qstart + i].label);
    double next_break = c->fh < p->pixel_y ? c->fh : p->pixel_y;
    double active_bottom = c->qlen > 0 ? c->q[c->qstart].bottom : DBL_MAX;
    if (active_bottom == p->sy1) {
        while (p->num_active > 0) {
            drec *t = &c->q[c->qstart];
            if ((t->h & D_PERSIST_AND_SLOPE) == D_PERSIST_BOTTOM) flip(c, t->label);
            c->qstart++; c->qlen--; p->num_active--;
        }
    }
    double next_bottom;
    if (p->num_active > 0) {
        next_bottom = active_bottom < next_break ? active_bottom : next_break;
    } else {
        double next_top = p->num_active < c->qlen ? c->q[c->qstart + p->num_active].top : DBL_MAX;
        if (next_top > p->sy1) {
            next_bottom = next_break < next_top ? next_break : next_top;
        } else {
            double s = split_next(c, p);
            next_bottom = next_break < s ? next_break : s;
            while (p->num_active > 0 && c->q[c->qstart].top == c->q[c->qstart].bottom) {
                drec *t = &c->q[c->qstart];
                if ((t->h & D_PERSIST_AND_SLOPE) == D_PERSIST_TOP) flip(c, t->label);
                c->qstart++; c->qlen--; p->num_active--;
            }
            for (int i = 0; i < p->num_active; i++) {
                drec *t = &c->q[c->qstart + i];
                if (t->top > 0.0 && (t->h & D_PERSIST_AND_SLOPE) == D_PERSIST_TOP) flip(c, t->label);
            }
        }
    }
    p->sy0 = p->sy1; p->sy1 = next_bottom;
    p->sx0 = p->sx1 = 0.0;
    p->current = 0;
}

/* Picture substances (readColor / getPicturePixel, Kernels.cl:856-860,
 * 1411-1442, and the build's bilinear filter): the texel colour at the
 * pixel's integer position (ax, ay), the sampling position computed in FP32
 * as the raster path computes it, the blend in FP64.  Texels outside the
 * picture are transparent. */
typedef struct {
    const float *subs;          /* raw substance records (4 floats each)   */
    const uint8_t *pict;        /* RGBA8 pile                              */
    const uint8_t *uses;        /* 24-byte PictureUse records              */
    int32_t n_uses;
    int32_t filter;             /* 0 nearest, 1 bilinear                   */
} dpict;

static void texel64(const dpict *pc, const uint8_t *use, long long x, long long y, double *out)
{
    int32_t w, h, off; memcpy(&w, use + 8, 4); memcpy(&h, use + 12, 4); memcpy(&off, use + 16, 4);
    if (x < 0 || y < 0 || x >= w || y >= h) { out[0] = out[1] = out[2] = out[3] = 0.0; return; }
    const uint8_t *px = pc->pict + off + 4 * (y * (long long)w + x);
    for (int i = 0; i < 4; i++) out[i] = (double)px[i] / 255.0;
}

static int f32_to_int(float v, int floor_mode)
{
    if (!(v == v)) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return floor_mode ? (int)floorf(v) : (int)v;
}

static void picture_color(const dpict *pc, uint64_t sid, int ax, int ay, double *out)
{
    uint32_t id; memcpy(&id, pc->subs + 4 * sid, 4);
    if ((int32_t)id < 0 || (int32_t)id >= pc->n_uses) { out[0] = out[1] = out[2] = out[3] = 0.0; return; }
    const uint8_t *use = pc->uses + 24L * id;
    float tx, ty, scale; memcpy(&tx, use, 4); memcpy(&ty, use + 4, 4); memcpy(&scale, use + 20, 4);
    scale = scale < 0.0000001f ? 0.0000001f : scale;
    const float sx = ((float)ax / scale) - tx, sy = ((float)ay / scale) - ty;
    if (pc->filter == 1) {
        const float fx = sx - 0.5f, fy = sy - 0.5f;
        const int ix = f32_to_int(fx, 1), iy = f32_to_int(fy, 1);
        const double wx = (double)(fx - (float)ix), wy = (double)(fy - (float)iy);
        double t0[4], t1[4], t2[4], t3[4];
        texel64(pc, use, ix, iy, t0); texel64(pc, use, (long long)ix + 1, iy, t1);
        texel64(pc, use, ix, (long long)iy + 1, t2); texel64(pc, use, (long long)ix + 1, (long long)iy + 1, t3);
        for (int i = 0; i < 4; i++)
            out[i] = ((t0[i] * (1 - wx)) + (t1[i] * wx)) * (1 - wy) + ((t2[i] * (1 - wx)) + (t3[i] * wx)) * wy;
        return;
    }
    texel64(pc, use, f32_to_int(sx, 0), f32_to_int(sy, 0), out);
}

static void section_color(const dcol *c, int n_shapes, const dshape *shapes, const double *rgba,
                          const double *bg, double *out, const dpict *pc, int ax, int ay)
{
    double P[3] = { 0, 0, 0 }, A = 0.0;
    int done = 0;
    /* determineColor's merge of consecutive shapes with one substance id
     * (Kernels.cl:1447-1513): a new id composites if its shape adds (or
     * continues); a repeated id is never composited again, and only the
     * deprecated continue compound toggles whether the merged run is set */
    uint64_t last_id = ~0ull;
    int last_is_continue = 1, last_is_set = 0;
    for (int w = 0; w < c->nwords && !done; w++) {
        uint64_t bits = c->stack[w];
        while (bits) {
            int b = __builtin_ctzll(bits); bits &= bits - 1;
            int s = 64 * w + b;
            if (s >= n_shapes) break;
            const uint64_t tag = shapes[s].tag, sid = shapes[s].sid;
            const int is_cont = (tag & D_TAG_COMPOUND) == D_TAG_CONTINUE;
            const int is_add = (tag & D_TAG_COMPOUND) == D_TAG_ADD;
            if (sid == last_id) {
                if (last_is_continue) {
                    if (!is_cont) { last_is_set = is_add; last_is_continue = 0; }
                    else last_is_set = !last_is_set;
                }
                continue;
            }
            last_id = sid;
            last_is_set = is_add || is_cont;
            if (!last_is_set) continue;
            const double *col = rgba + 4 * sid;
            double pcol[4];
            if (pc && (tag & D_TAG_TYPE) == D_TAG_PICTURE) {
                picture_color(pc, sid, ax, ay, pcol);
                col = pcol;
            }
            double t = 1.0 - A;
            for (int i = 0; i < 3; i++) P[i] += col[i] * col[3] * t;
            A += col[3] * t;
            if (A == 1.0) { done = 1; break; }
        }
    }
    if (!done) {
        double t = 1.0 - A;
        for (int i = 0; i < 3; i++) P[i] += bg[i] * bg[3] * t;
        A += bg[3] * t;
    }
    if (A > 0.0) { for (int i = 0; i < 3; i++) out[i] = P[i] / A; }
    else out[0] = out[1] = out[2] = 0.0;
    out[3] = A;
}

static int diff_loss(int32_t n_shapes, const uint64_t *tags, const int32_t *soff, const int32_t *ooff,
                     const double *pairs, const double *rgba, const double *bg, int32_t W, int32_t H,
                     const float *target, int32_t x0, int32_t y0, int32_t x1, int32_t y1, int32_t seg,
                     const int32_t *seg_map, const dpict *pc,
                     double *loss_out, double *image_out, uint64_t *signature)
{
    /* seg_map (W*H, may be NULL): the height of the column segment each
     * pixel belongs to (a power of two; the tile tree's 2^dd), segments
     * starting at multiples of it: each column runs from the segment that
     * holds row y0 through the one that holds row y1 - 1 (the loss counts
     * every row of those segments, the image only the window's).
     * Without a map: seg > 0 cuts the columns into segments of seg rows at
     * multiples of seg (the window must start on one and end on one or at
     * the canvas edge); seg == 0: one segment over the window's rows */
    if (!seg_map && (seg < 0 || (seg > 0 && (y0 % seg != 0 || (y1 % seg != 0 && y1 != H))))) return -3;
    if (seg_map) {
        for (int x = x0; x < x1; x++)
            for (int y = y0; y < y1; y++) {
                const int32_t h = seg_map[(long)y * W + x];
                if (h <= 0 || (h & (h - 1))) return -3;
            }
    }
    const int step = seg > 0 ? seg : (y1 - y0 > 0 ? y1 - y0 : 1);
    uint64_t sig = 0xcbf29ce484222325ull;
    dshape *shapes = calloc((size_t)(n_shapes > 0 ? n_shapes : 1), sizeof(dshape));
    for (int i = 0; i < n_shapes; i++) {
        build_shape(&shapes[i], pairs, ooff, soff[i], soff[i + 1], W, H, &sig);
        shapes[i].sid = tags[i] & D_TAG_ID;
        shapes[i].tag = tags[i];
    }
    for (int i = 0; i < n_shapes; i++)   /* gradients and nonzero fill are not modelled */
        if ((tags[i] & D_TAG_TYPE) == D_TAG_TYPE || (tags[i] & 0x0800000000000000ull) ||
            ((tags[i] & D_TAG_TYPE) == D_TAG_PICTURE && !pc)) {
            for (int k = 0; k < n_shapes; k++) { free(shapes[k].curves); free(shapes[k].strand_start); free(shapes[k].src); }
            free(shapes);
            return -4;
        }
    int nwords = (n_shapes + 63) / 64 + 1;
    dcol c;
    c.q = malloc(sizeof(drec) * D_QCAP);
    c.stack = malloc(sizeof(uint64_t) * nwords);
    c.nwords = nwords;
    double loss = 0.0;
    int rc = 0;
    int *cand = malloc(sizeof(int) * (size_t)(n_shapes > 0 ? n_shapes : 1));
    for (int x = x0; x < x1 && rc == 0; x++) {
    /* the shapes some strand of which can reach column x (checkInRange),
     * newest first, as the tile lists hold them */
    int n_cand = 0;
    for (int si = n_shapes - 1; si >= 0; si--) {
        const dshape *s = &shapes[si];
        if (s->kept && !(s->xmin - (double)x > 1.0 || !(s->xmax - (double)x > 0.0))) cand[n_cand++] = si;
    }
    for (int sy = seg_map ? y0 - y0 % seg_map[(long)y0 * W + x] : y0, sy1 = 0; sy < y1 && rc == 0; sy = sy1) {
        if (seg_map) {
            const int h = seg_map[(long)sy * W + x];
            sy1 = sy + h < H ? sy + h : H;
        } else {
            sy1 = sy + step < y1 ? sy + step : y1;
        }
        c.height_i = sy1 - sy; c.fh = (double)(sy1 - sy);
        const double dx = x, dy = sy;
        c.qstart = D_QCAP; c.qlen = 0; c.overflow = 0;
        memset(c.stack, 0, sizeof(uint64_t) * nwords);
        uint64_t skipped_y = 0;
        /* shapes newest first; a shape no strand of which reaches the column
         * (checkInRange for each) adds nothing and encloses nothing */
        for (int ci = 0; ci < n_cand; ci++) {
            const int si = cand[ci];
            dshape *s = &shapes[si];
            /* every curve point lies below the segment: every record starts
             * at or below its bottom (not added) and none crosses the left
             * border above its top.  Every point lies above the segment: no
             * record reaches into it, and the closed outlines cross the left
             * border above the top an even number of times, so their
             * enclosure flips cancel.  Either way the shape changes nothing;
             * how many shapes are skipped is a decision of the signature. */
            if (s->ymin - dy >= c.fh || s->ymax - dy < 0.0) { skipped_y++; continue; }
            int added = 0, enclosed_shape = 0;
            for (int k = 0; k < s->n_strands; k++) {
                const dq *cs = s->curves + s->strand_start[k];
                int m = s->strand_start[k + 1] - s->strand_start[k];
                int enclosed_strand = 0;
                double lx = cs[0].a.x - dx, rx = cs[m - 1].b.x - dx;
                if (lx <= 1.0 && rx > 0.0) {
                    dtrav l, r;
                    l.xpos = lx > 0.0 ? lx : 0.0;
                    r.xpos = rx < 1.0 ? rx : 1.0;
                    find_curve(&l, cs, m, dx, dy, 1);
                    find_curve(&r, cs, m, dx, dy, 0);
                    sig = fnv(sig, (uint64_t)((x << 20) ^ (si << 8) ^ (l.index << 4) ^ r.index));
                    spawn(&c, si, &l, &r, &added, &enclosed_strand, &sig);
                }
                enclosed_shape = enclosed_shape != enclosed_strand;
            }
            if (enclosed_shape) flip(&c, si);
        }
        sig = fnv(sig, skipped_y);
        if (c.overflow) { rc = -1; break; }
        /* stable insertion sort of the queue (front first) */
        for (int i = 1; i < c.qlen; i++) {
            drec v = c.q[c.qstart + i];
            int j = i - 1;
            while (j >= 0 && is_below(&c.q[c.qstart + j], &v)) { c.q[c.qstart + j + 1] = c.q[c.qstart + j]; j--; }
            c.q[c.qstart + j + 1] = v;
        }
        sig = fnv(sig, (uint64_t)c.qlen);
        for (int i = 0; i < c.qlen; i++)   /* the sort order */
            sig = fnv(sig, ((uint64_t)c.q[c.qstart + i].label << 20) ^ (uint64_t)(c.q[c.qstart + i].h >> 30));
        dsweep p; memset(&p, 0, sizeof(p));
        p.sx1 = 1.0;
        long guard = 1L << 24;
        g_split_sig = &sig;
        for (int yi = 0; yi < c.height_i && rc == 0; yi++) {
            p.pixel_y = (double)(yi + 1);
            double acc[4] = { 0, 0, 0, 0 }, area_sum = 0.0;
            int sections = 0;
            while ((p.sx1 < 1.0) || (p.sy1 < p.pixel_y)) {
                if (--guard < 0) { rc = -2; break; }
                vertical_advance(&c, &p);
                if (c.overflow) { rc = -1; break; }
                double nx = 1.0;
                if (p.current < p.num_active) {
                    const drec *t = &c.q[c.qstart + p.current];
                    double xm = isect_x(t, p.sy0 + 0.5 * (p.sy1 - p.sy0));
                    nx = (xm >= 1.0) ? 0.0 : (xm > 0.0 ? xm : 0.0);
                    /* which side of the clamps a section boundary is on */
                    sig = fnv(sig, (uint64_t)((xm >= 1.0) + 2 * (xm <= 0.0) + 4 * p.num_active));
                }
                p.sx0 = p.sx1; p.sx1 = nx;
                double col[4];
                section_color(&c, n_shapes, shapes, rgba, bg, col, pc, x, sy + yi);
                double area = (p.sx1 - p.sx0) * (p.sy1 - p.sy0);
                for (int i = 0; i < 4; i++) acc[i] += col[i] * area;
                area_sum += area;
                if (p.current < p.num_active) flip(&c, c.q[c.qstart + p.current].label);
                p.current++;
                sections++;
            }
            sig = fnv(sig, (uint64_t)sections);
            long pix = (long)(sy + yi) * W + x;
            const int in_window = sy + yi >= y0 && sy + yi < y1;
            for (int i = 0; i < 4; i++) {
                double v = acc[i] / area_sum;
                double e = v - (double)target[4 * pix + i];
                loss += e * e;
                if (image_out && in_window) image_out[4 * ((long)(sy - y0 + yi) * (x1 - x0) + (x - x0)) + i] = v;
            }
            p.sx0 = 0.0; p.sy0 = p.pixel_y;
        }
    }
    }
    free(cand);
    g_split_sig = NULL;
    for (int i = 0; i < n_shapes; i++) { free(shapes[i].curves); free(shapes[i].strand_start); free(shapes[i].src); }
    free(shapes); free(c.q); free(c.stack);
    *loss_out = loss;
    if (signature) *signature = sig;
    return rc;
}

int go_diff_loss(int32_t n_shapes, const uint64_t *tags, const int32_t *soff, const int32_t *ooff,
                 const double *pairs, const double *rgba, const double *bg, int32_t W, int32_t H,
                 const float *target, int32_t x0, int32_t y0, int32_t x1, int32_t y1, int32_t seg,
                 double *loss_out, double *image_out, uint64_t *signature)
{
    return diff_loss(n_shapes, tags, soff, ooff, pairs, rgba, bg, W, H, target, x0, y0, x1, y1, seg,
                     NULL, NULL, loss_out, image_out, signature);
}

int go_diff_loss_ex(int32_t n_shapes, const uint64_t *tags, const int32_t *soff, const int32_t *ooff,
                    const double *pairs, const double *rgba, const double *bg, int32_t W, int32_t H,
                    const float *target, int32_t x0, int32_t y0, int32_t x1, int32_t y1, int32_t seg,
                    const int32_t *seg_map, const float *substances, const uint8_t *pict_data,
                    const uint8_t *pict_uses, int32_t n_pict_uses, int32_t filter,
                    double *loss_out, double *image_out, uint64_t *signature)
{
    dpict pc = { substances, pict_data, pict_uses, n_pict_uses, filter };
    return diff_loss(n_shapes, tags, soff, ooff, pairs, rgba, bg, W, H, target, x0, y0, x1, y1, seg,
                     seg_map, substances ? &pc : NULL, loss_out, image_out, signature);
}


/* ------------------------------------------------------------------------ */
/* The adjoint, restated in FP64 (go_diff_grad): the same algorithm as the   */
/* HIP adjoint -- per generated line the boundary weights sum w and sum w u, */
/* then the line's end points, then the strand curve points, then the        */
/* outline points through the knob split -- so its gradients can be compared */
/* with the HIP path's term for term, and with finite differences.           */
/* ------------------------------------------------------------------------ */
typedef struct { int shape, curve, kind; double ta, tb, X; } dend;   /* kind 0 P(ta), 1 chord y at X, 2 y of P(ta) */
typedef struct { d2 a, b; dend ea, eb; double sw, swu, gay, gby; } dorig;

typedef struct {
    dorig *o; int n, cap;
    int shape;                 /* shape being generated */
    int lcur, rcur;            /* curves found by the two searches */
} gctx;

typedef struct { double y, ta, tb; int vertical, min_at_b; } dcut;

static dcut intersect_t(dtrav t)
{
    double ta = 0.0, tb = 1.0;
    if (!(t.left.x == t.ctl.x && t.left.y == t.ctl.y)) {
        for (int it = 0; it < D_MAX_BISECT; it++) {
            d2 lm = { 0.5 * t.left.x + 0.5 * t.ctl.x, 0.5 * t.left.y + 0.5 * t.ctl.y };
            d2 rm = { 0.5 * t.ctl.x + 0.5 * t.right.x, 0.5 * t.ctl.y + 0.5 * t.right.y };
            d2 on = { 0.5 * lm.x + 0.5 * rm.x, 0.5 * lm.y + 0.5 * rm.y };
            double flat = fabs(t.ctl.x - on.x) + fabs(t.ctl.y - on.y);
            if (!(flat > D_FLAT)) break;
            double tm = 0.5 * (ta + tb);
            if (t.xpos < on.x) { t.ctl = lm; t.right = on; tb = tm; } else { t.left = on; t.ctl = rm; ta = tm; }
        }
    }
    dcut c = { 0, ta, tb, t.left.x == t.right.x, t.right.y < t.left.y };
    c.y = c.vertical ? (t.left.y < t.right.y ? t.left.y : t.right.y)
                     : ((t.right.y - t.left.y) / (t.right.x - t.left.x)) * (t.xpos - t.left.x) + t.left.y;
    return c;
}

static void add_line_g(dcol *c, gctx *g, d2 l, d2 r, int label, int *added, int *enclosed, dend el, dend er)
{
    drec t;
    t.top = l.y < r.y ? l.y : r.y;
    t.bottom = l.y < r.y ? r.y : l.y;
    t.left = l.x; t.right = r.x; t.label = label;
    int pos = l.y <= r.y;
    int persistent = (l.x != r.x) && (l.x == 0.0);
    t.h = (pos ? D_POS : 0u) | (persistent ? D_PERSIST : 0u);
    int keep = persistent || (t.bottom - t.top) >= D_MINCROP;
    if (!keep) return;
    *enclosed = *enclosed || ((t.top <= 0.0) && (t.h & D_PERSIST_AND_SLOPE) == D_PERSIST_TOP)
                          || ((t.bottom <= 0.0) && (t.h & D_PERSIST_AND_SLOPE) == D_PERSIST_BOTTOM);
    if (!(t.top < c->fh && t.bottom > 0.0 && t.left < 1.0)) return;
    if (t.top <= 0.0) {
        double sx = isect_x(&t, 0.0);
        if (pos) { t.top = 0.0; t.left = sx; t.h &= ~D_PERSIST; }
        else { t.top = 0.0; t.right = sx; }
    }
    if (t.right <= 0.0) { *enclosed = 1; return; }
    *added = 1;
    if (g->n == g->cap) { g->cap = 2 * g->cap + 64; g->o = realloc(g->o, sizeof(dorig) * g->cap); }
    dorig o = { l, r, el, er, 0.0, 0.0, 0.0, 0.0 };
    t.origin = g->n;
    g->o[g->n++] = o;
    push_front(c, t);
}

static void spawn_g(dcol *c, gctx *g, int label, const dtrav *l, const dtrav *r, int *added, int *enclosed)
{
    double yl;
    dend e1 = { g->shape, g->lcur, 0, 0.0, 0.0, 0.0 };
    if (l->left.x >= 0.0) yl = l->left.y;
    else {
        dcut k = intersect_t(*l);
        yl = k.y;
        if (k.vertical) { e1.kind = 2; e1.ta = k.min_at_b ? k.tb : k.ta; }
        else { e1.kind = 1; e1.ta = k.ta; e1.tb = k.tb; e1.X = l->xpos; }
    }
    dend e2 = { g->shape, g->lcur, 0, 1.0, 0.0, 0.0 };
    dend e3 = { g->shape, g->rcur, 0, 0.0, 0.0, 0.0 };
    dend e4 = { g->shape, g->rcur, 0, 1.0, 0.0, 0.0 };
    int lw = 0, rw = 0;
    if (l->right.x < 1.0 && l->right.x > 0.0) {
        d2 a = { l->xpos, yl };
        add_line_g(c, g, a, l->right, label, added, enclosed, e1, e2); lw = 1;
    }
    double yr;
    if (r->right.x <= 1.0) yr = r->right.y;
    else {
        dcut k = intersect_t(*r);
        yr = k.y;
        if (k.vertical) { e4.kind = 2; e4.ta = k.min_at_b ? k.tb : k.ta; }
        else { e4.kind = 1; e4.ta = k.ta; e4.tb = k.tb; e4.X = r->xpos; }
    }
    if (r->left.x > 0.0 && r->left.x < 1.0 && l->index != r->index) {
        d2 b = { r->xpos, yr };
        add_line_g(c, g, r->left, b, label, added, enclosed, e3, e4); rw = 1;
    }
    if (l->right.x < r->left.x || (!lw && !rw)) {
        int blr = lw || l->left.x == l->right.x, brl = rw || r->left.x == r->right.x;
        d2 bl = blr ? l->right : (d2){ l->xpos, yl };
        d2 br = brl ? r->left : (d2){ r->xpos, yr };
        add_line_g(c, g, bl, br, label, added, enclosed, blr ? e2 : e1, brl ? e3 : e4);
    }
}

static d2 bez_at(const dq *q, double t)
{
    double a = (1 - t) * (1 - t), b = 2 * t * (1 - t), cc = t * t;
    d2 r = { a * q->a.x + b * q->c.x + cc * q->b.x, a * q->a.y + b * q->c.y + cc * q->b.y };
    return r;
}

static void chain_curve_d(const dshape *s, int curve, const double *pairs, d2 gl, d2 gc, d2 gr, double *dpairs)
{
    const dsrc *cs = &s->src[curve];
    d2 g0 = gl, g1 = gr;
    if (cs->flags & 4) { g0 = gr; g1 = gl; }
    const double *P0 = pairs + 2 * cs->p0, *C = pairs + 2 * cs->c, *P1 = pairs + 2 * cs->p1;
    d2 d0, dc, d1;
    double gt = 0.0, t = cs->t, u = 1.0 - t;
    d2 dbdt = { 2 * (u * (C[0] - P0[0]) + t * (P1[0] - C[0])), 2 * (u * (C[1] - P0[1]) + t * (P1[1] - C[1])) };
    if (cs->flags & 1) {
        d0 = (d2){ g0.x + u * gc.x + u * u * g1.x, g0.y + u * gc.y + u * u * g1.y };
        dc = (d2){ t * gc.x + 2 * t * u * g1.x, t * gc.y + 2 * t * u * g1.y };
        d1 = (d2){ t * t * g1.x, t * t * g1.y };
        gt = gc.x * (C[0] - P0[0]) + gc.y * (C[1] - P0[1]) + g1.x * dbdt.x + g1.y * dbdt.y;
    } else if (cs->flags & 2) {
        d0 = (d2){ u * u * g0.x, u * u * g0.y };
        dc = (d2){ 2 * t * u * g0.x + u * gc.x, 2 * t * u * g0.y + u * gc.y };
        d1 = (d2){ t * t * g0.x + t * gc.x + g1.x, t * t * g0.y + t * gc.y + g1.y };
        gt = g0.x * dbdt.x + g0.y * dbdt.y + gc.x * (P1[0] - C[0]) + gc.y * (P1[1] - C[1]);
    } else { d0 = g0; dc = gc; d1 = g1; }
    if (gt != 0.0) {
        double den = P0[0] - 2 * C[0] + P1[0];
        if (den != 0.0) {
            double kk = gt / (den * den);
            d0.x += kk * (P1[0] - C[0]);
            dc.x += kk * (P0[0] - P1[0]);
            d1.x += kk * (C[0] - P0[0]);
        }
    }
    dpairs[2 * cs->p0] += d0.x; dpairs[2 * cs->p0 + 1] += d0.y;
    dpairs[2 * cs->c] += dc.x;  dpairs[2 * cs->c + 1] += dc.y;
    dpairs[2 * cs->p1] += d1.x; dpairs[2 * cs->p1 + 1] += d1.y;
}

static void chain_point_d(const dshape *s, int curve, double t, d2 g, const double *pairs, double *dpairs)
{
    double a = (1 - t) * (1 - t), b = 2 * t * (1 - t), cc = t * t;
    chain_curve_d(s, curve, pairs, (d2){ a * g.x, a * g.y }, (d2){ b * g.x, b * g.y }, (d2){ cc * g.x, cc * g.y }, dpairs);
}

static void apply_end_d(const dshape *shapes, const dend *e, double dx, double dy, d2 g,
                        const double *pairs, double *dpairs)
{
    const dshape *s = &shapes[e->shape];
    if (e->kind == 0) { chain_point_d(s, e->curve, e->ta, g, pairs, dpairs); return; }
    if (e->kind == 2) { chain_point_d(s, e->curve, e->ta, (d2){ 0.0, g.y }, pairs, dpairs); return; }
    dq q = s->curves[e->curve];
    q.a.x -= dx; q.a.y -= dy; q.c.x -= dx; q.c.y -= dy; q.b.x -= dx; q.b.y -= dy;
    d2 pa = bez_at(&q, e->ta), pb = bez_at(&q, e->tb);
    double dxab = pb.x - pa.x, v = (e->X - pa.x) / dxab, sl = (pb.y - pa.y) / dxab;
    chain_point_d(s, e->curve, e->ta, (d2){ -g.y * sl * (1 - v), g.y * (1 - v) }, pairs, dpairs);
    chain_point_d(s, e->curve, e->tb, (d2){ -g.y * sl * v, g.y * v }, pairs, dpairs);
}

/* straight colour of a section and, given dL/dC, the gradients of its layers */
static void section_grad(const dcol *c, int n_shapes, const dshape *shapes, const double *rgba,
                         const double *bg, const double *C, const double *gc, double *drgba)
{
    double S[4] = { C[0] * C[3], C[1] * C[3], C[2] * C[3], C[3] };
    if (!(S[3] > 0.0)) return;
    double gP[3] = { gc[0] / S[3], gc[1] / S[3], gc[2] / S[3] };
    double gA = gc[3] - (gc[0] * C[0] + gc[1] * C[1] + gc[2] * C[2]) / S[3];
    double F[4] = { 0, 0, 0, 0 };
    /* the composited layers, front first: determineColor's merge walk
     * (Kernels.cl:1447-1513) as in section_color */
    int layers[4096], nl = 0;
    uint64_t last_id = ~0ull;
    int last_is_continue = 1, last_is_set = 0;
    for (int w = 0; w < c->nwords; w++) {
        uint64_t bits = c->stack[w];
        while (bits && nl < 4096) {
            int b = __builtin_ctzll(bits); bits &= bits - 1;
            const int sh = 64 * w + b;
            if (sh >= n_shapes) continue;
            const uint64_t tag = shapes[sh].tag, sid = shapes[sh].sid;
            const int is_cont = (tag & D_TAG_COMPOUND) == D_TAG_CONTINUE;
            const int is_add = (tag & D_TAG_COMPOUND) == D_TAG_ADD;
            if (sid == last_id) {
                if (last_is_continue) {
                    if (!is_cont) { last_is_set = is_add; last_is_continue = 0; }
                    else last_is_set = !last_is_set;
                }
                continue;
            }
            last_id = sid;
            last_is_set = is_add || is_cont;
            if (last_is_set) layers[nl++] = sh;
        }
    }
    for (int i = 0; i < nl; i++) {
        const double *col = rgba + 4 * shapes[layers[i]].sid;
        double T = 1.0 - F[3];
        double pm[4] = { col[0] * col[3], col[1] * col[3], col[2] * col[3], col[3] };
        double Fn[4];
        for (int j = 0; j < 4; j++) Fn[j] = F[j] + pm[j] * T;
        if (T > 0.0) {
            double V[4];
            if (col[3] < 1.0) for (int j = 0; j < 4; j++) V[j] = (S[j] - Fn[j]) / (1.0 - col[3]);
            else {
                double U[4] = { 0, 0, 0, 0 };
                int one = 0;
                for (int m = i + 1; m < nl && !one; m++) {
                    const double *c2 = rgba + 4 * shapes[layers[m]].sid;
                    double t2 = 1.0 - U[3];
                    for (int j = 0; j < 3; j++) U[j] += c2[j] * c2[3] * t2;
                    U[3] += c2[3] * t2;
                    if (U[3] == 1.0) one = 1;
                }
                if (!one) { double t2 = 1.0 - U[3]; for (int j = 0; j < 3; j++) U[j] += bg[j] * bg[3] * t2; U[3] += bg[3] * t2; }
                for (int j = 0; j < 4; j++) V[j] = T * U[j];
            }
            double *d = drgba + 4 * shapes[layers[i]].sid;
            double ta = T * col[3];
            d[0] += gP[0] * ta; d[1] += gP[1] * ta; d[2] += gP[2] * ta;
            d[3] += T * (gP[0] * col[0] + gP[1] * col[1] + gP[2] * col[2] + gA)
                  - (gP[0] * V[0] + gP[1] * V[1] + gP[2] * V[2] + gA * V[3]);
        }
        for (int j = 0; j < 4; j++) F[j] = Fn[j];
        if (F[3] == 1.0) break;
    }
}

int go_diff_grad(int32_t n_shapes, const uint64_t *tags, const int32_t *soff, const int32_t *ooff,
                 const double *pairs, const double *rgba, const double *bg, int32_t W, int32_t H,
                 const float *target, int32_t seg, double *loss_out, double *dpairs, double *drgba,
                 int32_t n_pairs, int32_t n_subs)
{
    uint64_t sig = 0;
    dshape *shapes = calloc((size_t)(n_shapes > 0 ? n_shapes : 1), sizeof(dshape));
    for (int i = 0; i < n_shapes; i++) {
        build_shape(&shapes[i], pairs, ooff, soff[i], soff[i + 1], W, H, &sig);
        shapes[i].sid = tags[i] & D_TAG_ID;
        shapes[i].tag = tags[i];
    }
    memset(dpairs, 0, sizeof(double) * 4 * (size_t)n_pairs);
    memset(drgba, 0, sizeof(double) * 4 * (size_t)n_subs);
    int nwords = (n_shapes + 63) / 64 + 1;
    dcol c;
    c.q = malloc(sizeof(drec) * D_QCAP);
    drec *qcopy = malloc(sizeof(drec) * D_QCAP);
    c.stack = malloc(sizeof(uint64_t) * nwords);
    uint64_t *stack0 = malloc(sizeof(uint64_t) * nwords);
    c.nwords = nwords;
    gctx g = { NULL, 0, 0, 0, 0, 0 };
    double loss = 0.0;
    double *G = malloc(sizeof(double) * 4 * (size_t)(seg > 0 ? seg : 1));
    int rc = 0;
    for (int x = 0; x < W && rc == 0; x++)
    for (int y0 = 0; y0 < H && rc == 0; y0 += seg) {
        const int y1 = y0 + seg < H ? y0 + seg : H;
        const double dx = x, dy = y0;
        c.height_i = y1 - y0; c.fh = (double)(y1 - y0);
        c.qstart = D_QCAP; c.qlen = 0; c.overflow = 0;
        memset(c.stack, 0, sizeof(uint64_t) * nwords);
        g.n = 0;
        for (int si = n_shapes - 1; si >= 0; si--) {
            dshape *s = &shapes[si];
            if (!s->kept) continue;
            int added = 0, enclosed_shape = 0;
            g.shape = si;
            for (int k = 0; k < s->n_strands; k++) {
                const int base = s->strand_start[k];
                const dq *cs = s->curves + base;
                int m = s->strand_start[k + 1] - base;
                int enclosed_strand = 0;
                double lx = cs[0].a.x - dx, rx = cs[m - 1].b.x - dx;
                if (lx <= 1.0 && rx > 0.0) {
                    dtrav l, r;
                    l.xpos = lx > 0.0 ? lx : 0.0;
                    r.xpos = rx < 1.0 ? rx : 1.0;
                    find_curve(&l, cs, m, dx, dy, 1);
                    find_curve(&r, cs, m, dx, dy, 0);
                    g.lcur = base + l.index; g.rcur = base + r.index;
                    spawn_g(&c, &g, si, &l, &r, &added, &enclosed_strand);
                }
                enclosed_shape = enclosed_shape != enclosed_strand;
            }
            if (enclosed_shape) flip(&c, si);
        }
        if (c.overflow) { rc = -1; break; }
        for (int i = 1; i < c.qlen; i++) {
            drec v = c.q[c.qstart + i];
            int j = i - 1;
            while (j >= 0 && is_below(&c.q[c.qstart + j], &v)) { c.q[c.qstart + j + 1] = c.q[c.qstart + j]; j--; }
            c.q[c.qstart + j + 1] = v;
        }
        const int qs0 = c.qstart, ql0 = c.qlen;
        memcpy(qcopy + qs0, c.q + qs0, sizeof(drec) * (size_t)ql0);
        memcpy(stack0, c.stack, sizeof(uint64_t) * nwords);
        for (int pass = 0; pass < 2 && rc == 0; pass++) {
            c.qstart = qs0; c.qlen = ql0;
            memcpy(c.q + qs0, qcopy + qs0, sizeof(drec) * (size_t)ql0);
            memcpy(c.stack, stack0, sizeof(uint64_t) * nwords);
            dsweep p; memset(&p, 0, sizeof(p));
            p.sx1 = 1.0;
            long guard = 1L << 24;
            for (int yi = 0; yi < c.height_i && rc == 0; yi++) {
                p.pixel_y = (double)(yi + 1);
                double acc[4] = { 0, 0, 0, 0 }, area_sum = 0.0;
                int have_prev = 0, prev_inside = 0, prev_origin = 0;
                double prev_col[4] = { 0, 0, 0, 0 };
                long pix = (long)(y0 + yi) * W + x;
                /* span boundaries inside the row: sum C (e_j - e_{j-1}) at the
                 * span's top and bottom; a boundary set by a line end point
                 * moves with it, moving area between the two spans */
                double top_sum[4] = { 0, 0, 0, 0 }, bot_sum[4] = { 0, 0, 0, 0 };
                double e_top = 0.0, e_bot = 0.0;
                int span_open = 0, pend = 0, pend_org = -1, pend_end = 0, cur_org = -1, cur_end = 0;
                double pend_bot[4] = { 0, 0, 0, 0 };
                while ((p.sx1 < 1.0) || (p.sy1 < p.pixel_y)) {
                    if (--guard < 0) { rc = -2; break; }
                    vertical_advance(&c, &p);
                    if (c.overflow) { rc = -1; break; }
                    if (p.current == 0) {
                        have_prev = 0;
                        if (pass == 1) {
                            if (span_open) {
                                if (pend && pend_org >= 0) {
                                    double J = 0.0;
                                    for (int i = 0; i < 4; i++) J += G[4 * yi + i] * (pend_bot[i] - top_sum[i]);
                                    if (pend_end == 0) g.o[pend_org].gay += J; else g.o[pend_org].gby += J;
                                }
                                pend = 1; pend_org = cur_org; pend_end = cur_end;
                                memcpy(pend_bot, bot_sum, sizeof(bot_sum));
                            }
                            span_open = 1;
                            memset(top_sum, 0, sizeof(top_sum)); memset(bot_sum, 0, sizeof(bot_sum));
                            e_top = e_bot = 0.0;
                            cur_org = -1;
                            if (p.sy1 != p.pixel_y && p.sy1 != c.fh) {
                                for (int i = 0; i < c.qlen && cur_org < 0; i++) {
                                    const drec *t = &c.q[c.qstart + i];
                                    double yy = i < p.num_active ? t->bottom : t->top;
                                    if (i >= p.num_active && t->top > p.sy1) break;
                                    if (yy != p.sy1) continue;
                                    const dorig *o = &g.o[t->origin];
                                    if (o->a.y == p.sy1) { cur_org = t->origin; cur_end = 0; }
                                    else if (o->b.y == p.sy1) { cur_org = t->origin; cur_end = 1; }
                                }
                            }
                        }
                    }
                    double nx = 1.0, xm = 1.0;
                    if (p.current < p.num_active) {
                        const drec *t = &c.q[c.qstart + p.current];
                        xm = isect_x(t, p.sy0 + 0.5 * (p.sy1 - p.sy0));
                        nx = (xm >= 1.0) ? 0.0 : (xm > 0.0 ? xm : 0.0);
                    }
                    p.sx0 = p.sx1; p.sx1 = nx;
                    double col[4];
                    section_color(&c, n_shapes, shapes, rgba, bg, col, NULL, 0, 0);
                    double ddy = p.sy1 - p.sy0;
                    double area = (p.sx1 - p.sx0) * ddy;
                    if (pass == 1) {
                        double et = 1.0, eb = 1.0;
                        if (p.current < p.num_active) {
                            const drec *t = &c.q[c.qstart + p.current];
                            if (xm > 0.0 && xm < 1.0) { et = isect_x(t, p.sy0); eb = isect_x(t, p.sy1); }
                            else et = eb = nx;
                        }
                        for (int i = 0; i < 4; i++) { top_sum[i] += col[i] * (et - e_top); bot_sum[i] += col[i] * (eb - e_bot); }
                        e_top = et; e_bot = eb;
                    }
                    if (pass == 0) {
                        for (int i = 0; i < 4; i++) acc[i] += col[i] * area;
                        area_sum += area;
                    } else {
                        if (have_prev && prev_inside && ddy > 0.0) {
                            double w = 0.0;
                            for (int i = 0; i < 4; i++) w += G[4 * yi + i] * (prev_col[i] - col[i]);
                            w *= ddy;
                            dorig *o = &g.o[prev_origin];
                            double ym = p.sy0 + 0.5 * ddy;
                            double u = (o->b.y != o->a.y) ? (ym - o->a.y) / (o->b.y - o->a.y) : 0.0;
                            o->sw += w; o->swu += w * u;
                        }
                        double gcol[4];
                        for (int i = 0; i < 4; i++) gcol[i] = G[4 * yi + i] * area;
                        section_grad(&c, n_shapes, shapes, rgba, bg, col, gcol, drgba);
                    }
                    if (p.current < p.num_active) {
                        const drec *t = &c.q[c.qstart + p.current];
                        prev_origin = t->origin; prev_inside = xm > 0.0 && xm < 1.0;
                        memcpy(prev_col, col, sizeof(col)); have_prev = 1;
                        flip(&c, t->label);
                    } else have_prev = 0;
                    p.current++;
                }
                if (pass == 1 && span_open && pend && pend_org >= 0) {
                    double J = 0.0;
                    for (int i = 0; i < 4; i++) J += G[4 * yi + i] * (pend_bot[i] - top_sum[i]);
                    if (pend_end == 0) g.o[pend_org].gay += J; else g.o[pend_org].gby += J;
                }
                if (pass == 0) {
                    for (int i = 0; i < 4; i++) {
                        double v = acc[i] / area_sum, e = v - (double)target[4 * pix + i];
                        loss += e * e;
                        G[4 * yi + i] = 2.0 * e / area_sum;
                    }
                }
                p.sx0 = 0.0; p.sy0 = p.pixel_y;
            }
        }
        /* line weights -> end points -> outline points */
        for (int i = 0; i < g.n; i++) {
            const dorig *o = &g.o[i];
            if (o->sw == 0.0 && o->swu == 0.0 && o->gay == 0.0 && o->gby == 0.0) continue;
            double gax = o->sw - o->swu, gbx = o->swu;
            double kk = (o->b.y != o->a.y) ? (o->b.x - o->a.x) / (o->b.y - o->a.y) : 0.0;
            apply_end_d(shapes, &o->ea, dx, dy, (d2){ gax, -kk * gax + o->gay }, pairs, dpairs);
            apply_end_d(shapes, &o->eb, dx, dy, (d2){ gbx, -kk * gbx + o->gby }, pairs, dpairs);
        }
    }
    for (int i = 0; i < n_shapes; i++) { free(shapes[i].curves); free(shapes[i].strand_start); free(shapes[i].src); }
    free(shapes); free(c.q); free(qcopy); free(c.stack); free(stack0); free(g.o); free(G);
    *loss_out = loss;
    return rc;
}
