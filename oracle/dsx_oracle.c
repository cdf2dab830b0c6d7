/*
 * dsx_oracle.c -- CPU restatement of datashader 0.19.1's points / line aggregation.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is compared
 * against; it is never part of the product.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library (libdsx.so) never
 * links or calls it.
 *
 * It restates, in sequential row order, what the reference's numba kernels compute:
 *   - point projection       glyphs/points.py:188-212 (A.1 of the spec)
 *   - per-base append        reductions.py:554-558 (count(v)), 582-584 (count()),
 *                            1056-1063 (sum), 958-962 (_sum_zero), 1128-1137 (m2),
 *                            1178-1182 (min), 1223-1227 (max), 842-846 / 859-861 (any),
 *                            and the generated append order of compiler.py:321-475
 *                            (temps read before any base is updated, 453-454)
 *   - categorical slicing    compiler.py:379-390 (agg[:, :, int(code)]) with numba's
 *                            negative-index wraparound for code -1
 *   - line drawing           glyphs/line.py:688-720 (map_onto_pixel_snap),
 *                            734-800 (_liang_barsky, _clipt), 986-1031 (_bresenham),
 *                            1033-1097 (draw_segment), 1099-1136 (axis-0 extend),
 *                            1340-1386 (axis-1 x-constant extend)
 *   - bounds                 glyphs/glyph.py:66-78 (_compute_bounds_numba)
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math (no FMA contraction, IEEE division),
 * which is what numba's non-fastmath LLVM code computes.  Values are passed as float64:
 * the reference promotes float32 columns to float64 before every projection and
 * accumulation, so promoting first is exact and equivalent.  The only float32
 * arithmetic the reference performs on this path is the segment delta in
 * _liang_barsky (dx1 = x1 - x0 typed by the column dtype, line.py:752,758); the
 * xf32 / yf32 flags reproduce it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    /* canvas */
    int32_t W, H, ncat;
    /* per-row inputs (NULL when absent) */
    const double *val;   /* value column promoted to float64 */
    const int64_t *cat;  /* category codes */
    /* outputs, each [H][W][ncat] (ncat = 1 when not categorical), NULL when absent */
    uint32_t *count;     /* count()            u32, 0-init   */
    uint32_t *countv;    /* count(v)           u32, 0-init   */
    uint8_t *any;        /* any()              bool, 0-init  */
    uint8_t *anyv;       /* any(v)             bool, 0-init  */
    double *sum;         /* sum(v)  CPU base   f64, NaN-init */
    double *sumz;        /* _sum_zero(v)       f64, 0-init   */
    double *min;         /* min(v)             f64, NaN-init */
    double *max;         /* max(v)             f64, NaN-init */
    double *m2;          /* m2(v)              f64, 0-init   */
    double *m2_sum;      /* m2 temp: _sum_zero(v), 0-init    */
    uint32_t *m2_cnt;    /* m2 temp: count(v), 0-init        */
} dsx_oracle_aggs;

typedef struct {
    double sx, tx, sy, ty;
    double xmin, xmax, ymin, ymax;
    int64_t xxmax, yymax;  /* lines: round(xmax*sx+tx), round(ymax*sy+ty) (line.py:716-717) */
    int32_t xf32, yf32;    /* lines: coordinate columns are float32 (segment delta in f32) */
} dsx_oracle_view;

/* reductions.py append bodies, applied in compiler.py's order for row i at (x, y). */
static void oracle_append(const dsx_oracle_aggs *a, int64_t i, int64_t x, int64_t y)
{
    int64_t c = 0;
    if (a->cat) {
        c = a->cat[i];
        if (c < 0) c += a->ncat;          /* numba wraparound of agg[:, :, -1] */
        if (c < 0 || c >= a->ncat) return; /* outside the aggregate: no defined result */
    }
    /* numba wraps negative indices of agg[y, x] as well; an index still outside the
     * aggregate would be an out-of-bounds write in the reference: dropped here */
    if (x < 0) x += a->W;
    if (y < 0) y += a->H;
    if (x < 0 || x >= a->W || y < 0 || y >= a->H) return;
    const int64_t cell = (y * (int64_t)a->W + x) * (int64_t)a->ncat + c;
    const double v = a->val ? a->val[i] : 0.0;
    const int vnan = a->val ? isnan(v) : 0;

    /* temps read before any update (compiler.py:453-454) */
    double s0 = 0.0;
    uint32_t n0 = 0;
    if (a->m2) { s0 = a->m2_sum[cell]; n0 = a->m2_cnt[cell]; }

    if (a->count) a->count[cell] += 1u;                                   /* 582-584 */
    if (a->countv && !vnan) a->countv[cell] += 1u;                        /* 554-558 */
    if (a->any) a->any[cell] = 1;                                         /* 859-861 */
    if (a->anyv && !vnan) a->anyv[cell] = 1;                              /* 842-846 */
    if (a->sum && !vnan) {                                                /* 1056-1063 */
        if (isnan(a->sum[cell])) a->sum[cell] = v;
        else a->sum[cell] += v;
    }
    if (a->sumz && !vnan) a->sumz[cell] += v;                             /* 958-962 */
    if (a->min && !vnan && (isnan(a->min[cell]) || a->min[cell] > v))     /* 1178-1182 */
        a->min[cell] = v;
    if (a->max && !vnan && (isnan(a->max[cell]) || a->max[cell] < v))     /* 1223-1227 */
        a->max[cell] = v;
    if (a->m2) {                                                          /* 1128-1137 */
        if (!vnan) {
            if (n0 > 0) {
                const double u1 = s0 / (double)n0;
                const double u = (s0 + v) / (double)((int64_t)n0 + 1);
                a->m2[cell] += (v - u1) * (v - u);
            }
            a->m2_sum[cell] += v;
            a->m2_cnt[cell] += 1u;
        }
    }
}

/* ---- points: glyphs/points.py:188-212 ---------------------------------------- */

/* Bin index of every row (y*W + x), or -1 when the row is dropped. */
void dsx_oracle_point_bins(const double *xs, const double *ys, int64_t n,
                           const dsx_oracle_view *v, int32_t W, int32_t H, int64_t *out)
{
    for (int64_t i = 0; i < n; ++i) {
        const double x = xs[i], y = ys[i];
        if ((v->xmin <= x && x <= v->xmax) && (v->ymin <= y && y <= v->ymax)) {
            int64_t xx = (int64_t)(x * v->sx + v->tx);
            int64_t yy = (int64_t)(y * v->sy + v->ty);
            int64_t xi = xx >= W ? W - 1 : xx;
            int64_t yi = yy >= H ? H - 1 : yy;
            if (xi < 0) xi += W;
            if (yi < 0) yi += H;
            out[i] = yi * (int64_t)W + xi;
        } else {
            out[i] = -1;
        }
    }
}

void dsx_oracle_points(const double *xs, const double *ys, int64_t n,
                       const dsx_oracle_view *v, const dsx_oracle_aggs *a)
{
    const int64_t W = a->W, H = a->H;
    for (int64_t i = 0; i < n; ++i) {
        const double x = xs[i], y = ys[i];
        if ((v->xmin <= x && x <= v->xmax) && (v->ymin <= y && y <= v->ymax)) {
            const int64_t xx = (int64_t)(x * v->sx + v->tx);
            const int64_t yy = (int64_t)(y * v->sy + v->ty);
            oracle_append(a, i, xx >= W ? W - 1 : xx, yy >= H ? H - 1 : yy);
        }
    }
}

/* ---- lines: glyphs/line.py ----------------------------------------------------- */

/* _clipt, line.py:783-800 */
static int clipt(double p, double q, double *t0, double *t1)
{
    int accept = 1;
    if (p < 0 && q < 0) {
        const double r = q / p;
        if (r > *t1) accept = 0;
        else if (r > *t0) *t0 = r;
    } else if (p > 0 && q < p) {
        const double r = q / p;
        if (r < *t0) accept = 0;
        else if (r < *t1) *t1 = r;
    } else if (q < 0) {
        accept = 0;
    }
    return accept;
}

/* map_onto_pixel_snap, line.py:688-720 */
static void snap(const dsx_oracle_view *v, double x, double y, int64_t *px, int64_t *py)
{
    const int64_t xx = (int64_t)(x * v->sx + v->tx);
    const int64_t yy = (int64_t)(y * v->sy + v->ty);
    *px = xx == v->xxmax ? xx - 1 : xx;
    *py = yy == v->yymax ? yy - 1 : yy;
}

/* _bresenham, line.py:986-1031 */
static void bresenham(const dsx_oracle_aggs *a, int64_t i, int segment_start,
                      int64_t x0, int64_t x1, int64_t y0, int64_t y1, int clipped)
{
    int64_t dx = x1 - x0;
    const int64_t ix = (dx > 0) - (dx < 0);
    dx = (dx < 0 ? -dx : dx) * 2;
    int64_t dy = y1 - y0;
    const int64_t iy = (dy > 0) - (dy < 0);
    dy = (dy < 0 ? -dy : dy) * 2;

    if (!clipped && !(dx | dy)) {
        oracle_append(a, i, x0, y0);
        return;
    }
    if (segment_start) oracle_append(a, i, x0, y0);

    if (dx >= dy) {
        int64_t error = 2 * dy - dx;
        while (x0 != x1) {
            if (error >= 0 && (error || ix > 0)) {
                error -= 2 * dx;
                y0 += iy;
            }
            error += 2 * dy;
            x0 += ix;
            oracle_append(a, i, x0, y0);
        }
    } else {
        int64_t error = 2 * dx - dy;
        while (y0 != y1) {
            if (error >= 0 && (error || iy > 0)) {
                error -= 2 * dy;
                x0 += ix;
            }
            error += 2 * dx;
            y0 += iy;
            oracle_append(a, i, x0, y0);
        }
    }
}

/* draw_segment (line_width == 0), line.py:1033-1097 with _liang_barsky 734-780 */
static void draw_segment(const dsx_oracle_aggs *a, const dsx_oracle_view *v, int64_t i,
                         int segment_start, double x0, double x1, double y0, double y1)
{
    int skip = isnan(x0) || isnan(y0) || isnan(x1) || isnan(y1);

    if (x0 < v->xmin && x1 < v->xmin) skip = 1;
    else if (x0 > v->xmax && x1 > v->xmax) skip = 1;
    else if (y0 < v->ymin && y1 < v->ymin) skip = 1;
    else if (y0 > v->ymax && y1 > v->ymax) skip = 1;

    double t0 = 0.0, t1 = 1.0;
    const double dx1 = v->xf32 ? (double)((float)x1 - (float)x0) : x1 - x0;
    if (!clipt(-dx1, x0 - v->xmin, &t0, &t1)) skip = 1;
    if (!clipt(dx1, v->xmax - x0, &t0, &t1)) skip = 1;
    const double dy1 = v->yf32 ? (double)((float)y1 - (float)y0) : y1 - y0;
    if (!clipt(-dy1, y0 - v->ymin, &t0, &t1)) skip = 1;
    if (!clipt(dy1, v->ymax - y0, &t0, &t1)) skip = 1;

    int clipped_end = 0, clipped_start = 0;
    if (t1 < 1) {
        clipped_end = 1;
        x1 = x0 + t1 * dx1;
        y1 = y0 + t1 * dy1;
    }
    if (t0 > 0) {
        clipped_start = 1;
        x0 = x0 + t0 * dx1;
        y0 = y0 + t0 * dy1;
    }
    if (skip) return;

    const int clipped = clipped_start || clipped_end;
    segment_start = segment_start || clipped_start;
    int64_t px0, py0, px1, py1;
    snap(v, x0, y0, &px0, &py0);
    snap(v, x1, y1, &px1, &py1);
    bresenham(a, i, segment_start, px0, px1, py0, py1, clipped);
}

/* Single segment, as the reference's test helper drives draw_segment (test_glyphs.py:73-82). */
void dsx_oracle_draw_segment(const dsx_oracle_view *v, const dsx_oracle_aggs *a, int64_t i,
                             int32_t segment_start, double x0, double y0, double x1, double y1)
{
    draw_segment(a, v, i, segment_start, x0, x1, y0, y1);
}

/* LineAxis0: one line through all rows, line.py:1099-1136. */
void dsx_oracle_line_axis0(const double *xs, const double *ys, int64_t n, int32_t plot_start,
                           const dsx_oracle_view *v, const dsx_oracle_aggs *a)
{
    for (int64_t i = 0; i + 1 < n; ++i) {
        const int segment_start =
            i == 0 ? plot_start : (isnan(xs[i - 1]) || isnan(ys[i - 1]));
        draw_segment(a, v, i, segment_start, xs[i], xs[i + 1], ys[i], ys[i + 1]);
    }
}

/* LinesAxis1XConstant: one line per row, shared x, ys row-major [nrows][ncols];
 * line.py:1340-1386. */
void dsx_oracle_lines_xconst(const double *xs, int64_t ncols, const double *ys, int64_t nrows,
                             const dsx_oracle_view *v, const dsx_oracle_aggs *a)
{
    for (int64_t i = 0; i < nrows; ++i) {
        const double *row = ys + i * ncols;
        for (int64_t j = 0; j + 1 < ncols; ++j) {
            const int segment_start = (j == 0) || isnan(xs[j - 1]) || isnan(row[j - 1]);
            draw_segment(a, v, i, segment_start, xs[j], xs[j + 1], row[j], row[j + 1]);
        }
    }
}

/* LinesAxis1: one line per row, per-row x and y, both row-major [nrows][ncols];
 * line.py:1244-1336 (_build_extend_line_axis1_none_constant). */
void dsx_oracle_lines_axis1(const double *xs, const double *ys, int64_t nrows, int64_t ncols,
                            const dsx_oracle_view *v, const dsx_oracle_aggs *a)
{
    for (int64_t i = 0; i < nrows; ++i) {
        const double *xr = xs + i * ncols, *yr = ys + i * ncols;
        for (int64_t j = 0; j + 1 < ncols; ++j) {
            const int segment_start = (j == 0) || isnan(xr[j - 1]) || isnan(yr[j - 1]);
            draw_segment(a, v, i, segment_start, xr[j], xr[j + 1], yr[j], yr[j + 1]);
        }
    }
}

/* ---- bounds: glyph.py:66-78 ---------------------------------------------------- */
void dsx_oracle_bounds(const double *arr, int64_t n, double *out2)
{
    double minval = INFINITY, maxval = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        const double x = arr[i];
        if (!isnan(x)) {
            if (x < minval) minval = x;
            if (x > maxval) maxval = x;
        }
    }
    out2[0] = minval;
    out2[1] = maxval;
}
