/*
 * cpu_baseline.c -- the CPU baseline the GPU numbers are reported beside.
 *
 * TEST INFRASTRUCTURE ONLY: it is timed by bench.py's cpu_baseline leg and checked
 * against the sequential oracle (dsx_oracle.c) by tests/test_oracle.py; the product
 * library never links or calls it.
 *
 * Two forms of the reference's CPU path, specialised per workload (no per-row branch on
 * which planes exist) and built -O3:
 *   - one thread: the pandas path (data_libraries/pandas.py:26-66), whose numba kernels
 *     are ngjit without parallel=True (utils.py:52), i.e. one core;
 *   - P threads: the dask path with npartitions = P (data_libraries/dask.py:168-217):
 *     each thread aggregates one contiguous partition into a private canvas (`chunk`:
 *     create + extend), the canvases are merged by a binary tree (`combine` with
 *     split_every=2, each merge one task), then finalize runs once.
 * Projection and appends follow glyphs/points.py:188-212 and reductions.py (count
 * 582-584, count(v) 554-558, _sum_zero 958-962); lines follow glyphs/line.py:688-720,
 * 734-800, 986-1097, 1340-1386 (LinesAxis1XConstant, float32 y columns).
 * Built with -ffp-contract=off (no FMA contraction), like the oracle.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double sx, tx, sy, ty;
    double xmin, xmax, ymin, ymax;
    int64_t xxmax, yymax;
} cpub_view;

/* ---- points kernels: one partition [lo, hi) into one canvas ------------------------ */

static void count_mean_f32(const float *xs, const float *ys, const float *vs, int64_t lo,
                           int64_t hi, const cpub_view *v, int32_t W, int32_t H,
                           uint32_t *count, uint32_t *countv, double *sum)
{
    for (int64_t i = lo; i < hi; ++i) {
        const double x = xs[i], y = ys[i];
        if ((v->xmin <= x && x <= v->xmax) && (v->ymin <= y && y <= v->ymax)) {
            const int64_t xx = (int64_t)(x * v->sx + v->tx);
            const int64_t yy = (int64_t)(y * v->sy + v->ty);
            const int64_t xi = xx >= W ? W - 1 : xx, yi = yy >= H ? H - 1 : yy;
            const int64_t c = yi * W + xi;
            const double val = vs[i];
            count[c] += 1u;
            if (!isnan(val)) {
                countv[c] += 1u;
                sum[c] += val;
            }
        }
    }
}

static void count_f64(const double *xs, const double *ys, int64_t lo, int64_t hi,
                      const cpub_view *v, int32_t W, int32_t H, uint32_t *count)
{
    for (int64_t i = lo; i < hi; ++i) {
        const double x = xs[i], y = ys[i];
        if ((v->xmin <= x && x <= v->xmax) && (v->ymin <= y && y <= v->ymax)) {
            const int64_t xx = (int64_t)(x * v->sx + v->tx);
            const int64_t yy = (int64_t)(y * v->sy + v->ty);
            const int64_t xi = xx >= W ? W - 1 : xx, yi = yy >= H ? H - 1 : yy;
            count[yi * W + xi] += 1u;
        }
    }
}

static void count_cat_f32(const float *xs, const float *ys, const uint8_t *cat, int64_t lo,
                          int64_t hi, const cpub_view *v, int32_t W, int32_t H, int32_t ncat,
                          uint32_t *count)
{
    for (int64_t i = lo; i < hi; ++i) {
        const double x = xs[i], y = ys[i];
        if ((v->xmin <= x && x <= v->xmax) && (v->ymin <= y && y <= v->ymax)) {
            const int64_t xx = (int64_t)(x * v->sx + v->tx);
            const int64_t yy = (int64_t)(y * v->sy + v->ty);
            const int64_t xi = xx >= W ? W - 1 : xx, yi = yy >= H ? H - 1 : yy;
            const int64_t c = cat[i];
            if (c < ncat) count[(yi * W + xi) * ncat + c] += 1u;
        }
    }
}

/* ---- lines: LinesAxis1XConstant, count ---------------------------------------------- */

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

static inline void put(uint32_t *count, int32_t W, int32_t H, int64_t x, int64_t y)
{
    if (x < 0) x += W;
    if (y < 0) y += H;
    if (x >= 0 && x < W && y >= 0 && y < H) count[y * W + x] += 1u;
}

/* draw_segment (line_width 0) with float64 x and float32 y: the y delta is float32 */
static void segment(const cpub_view *v, int32_t W, int32_t H, uint32_t *count,
                    int segment_start, double x0, double x1, float fy0, float fy1)
{
    double y0 = fy0, y1 = fy1;
    int skip = isnan(x0) || isnan(y0) || isnan(x1) || isnan(y1);
    if (x0 < v->xmin && x1 < v->xmin) skip = 1;
    else if (x0 > v->xmax && x1 > v->xmax) skip = 1;
    else if (y0 < v->ymin && y1 < v->ymin) skip = 1;
    else if (y0 > v->ymax && y1 > v->ymax) skip = 1;
    double t0 = 0.0, t1 = 1.0;
    const double dx1 = x1 - x0;
    if (!clipt(-dx1, x0 - v->xmin, &t0, &t1)) skip = 1;
    if (!clipt(dx1, v->xmax - x0, &t0, &t1)) skip = 1;
    const double dy1 = (double)(fy1 - fy0);
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
    int64_t px0 = (int64_t)(x0 * v->sx + v->tx), py0 = (int64_t)(y0 * v->sy + v->ty);
    int64_t px1 = (int64_t)(x1 * v->sx + v->tx), py1 = (int64_t)(y1 * v->sy + v->ty);
    if (px0 == v->xxmax) px0 -= 1;
    if (py0 == v->yymax) py0 -= 1;
    if (px1 == v->xxmax) px1 -= 1;
    if (py1 == v->yymax) py1 -= 1;
    int64_t dx = px1 - px0;
    const int64_t ix = (dx > 0) - (dx < 0);
    dx = (dx < 0 ? -dx : dx) * 2;
    int64_t dy = py1 - py0;
    const int64_t iy = (dy > 0) - (dy < 0);
    dy = (dy < 0 ? -dy : dy) * 2;
    if (!clipped && !(dx | dy)) {
        put(count, W, H, px0, py0);
        return;
    }
    if (segment_start) put(count, W, H, px0, py0);
    if (dx >= dy) {
        int64_t error = 2 * dy - dx;
        while (px0 != px1) {
            if (error >= 0 && (error || ix > 0)) {
                error -= 2 * dx;
                py0 += iy;
            }
            error += 2 * dy;
            px0 += ix;
            put(count, W, H, px0, py0);
        }
    } else {
        int64_t error = 2 * dx - dy;
        while (py0 != py1) {
            if (error >= 0 && (error || iy > 0)) {
                error -= 2 * dy;
                px0 += ix;
            }
            error += 2 * dx;
            py0 += iy;
            put(count, W, H, px0, py0);
        }
    }
}

static void lines_xconst_f32(const double *xs, int64_t ncols, const float *ys, int64_t lo,
                             int64_t hi, const cpub_view *v, int32_t W, int32_t H,
                             uint32_t *count)
{
    for (int64_t i = lo; i < hi; ++i) {
        const float *row = ys + i * ncols;
        for (int64_t j = 0; j + 1 < ncols; ++j) {
            const int start = (j == 0) || isnan(xs[j - 1]) || isnan(row[j - 1]);
            segment(v, W, H, count, start, xs[j], xs[j + 1], row[j], row[j + 1]);
        }
    }
}

/* ---- partitions, private canvases and the binary combine tree ------------------------ */

enum { K_COUNT_MEAN = 1, K_COUNT_F64 = 2, K_COUNT_CAT = 3, K_LINES = 4 };

typedef struct {
    int kind;
    const void *x, *y, *v;   /* columns (lines: x = shared x, y = [rows][ncols])       */
    int64_t n, ncols;        /* rows (lines: series)                                    */
    cpub_view view;
    int32_t W, H, ncat;
    int64_t ncells;
    int nparts;
    uint32_t **count, **countv;  /* per partition (partition 0's are the outputs)       */
    double **sum;
} job;

typedef struct {
    job *j;
    int part;
    int stride;              /* combine: merge partition part + stride into part        */
} task;

static void *run_chunk(void *arg)
{
    task *t = (task *)arg;
    job *j = t->j;
    const int64_t lo = j->n * t->part / j->nparts, hi = j->n * (t->part + 1) / j->nparts;
    const int p = t->part;
    const size_t cb = (size_t)j->ncells;
    /* create: zero-initialised planes (reductions.py:446-472) */
    memset(j->count[p], 0, cb * 4);
    if (j->countv) memset(j->countv[p], 0, cb * 4);
    if (j->sum) memset(j->sum[p], 0, cb * 8);
    switch (j->kind) {
    case K_COUNT_MEAN:
        count_mean_f32(j->x, j->y, j->v, lo, hi, &j->view, j->W, j->H, j->count[p],
                       j->countv[p], j->sum[p]);
        break;
    case K_COUNT_F64:
        count_f64(j->x, j->y, lo, hi, &j->view, j->W, j->H, j->count[p]);
        break;
    case K_COUNT_CAT:
        count_cat_f32(j->x, j->y, j->v, lo, hi, &j->view, j->W, j->H, j->ncat, j->count[p]);
        break;
    default:
        lines_xconst_f32(j->x, j->ncols, j->y, lo, hi, &j->view, j->W, j->H, j->count[p]);
        break;
    }
    return NULL;
}

static void *run_combine(void *arg)
{
    task *t = (task *)arg;
    job *j = t->j;
    const int a = t->part, b = t->part + t->stride;
    for (int64_t c = 0; c < j->ncells; ++c) j->count[a][c] += j->count[b][c];
    if (j->countv)
        for (int64_t c = 0; c < j->ncells; ++c) j->countv[a][c] += j->countv[b][c];
    if (j->sum)
        for (int64_t c = 0; c < j->ncells; ++c) j->sum[a][c] += j->sum[b][c];
    return NULL;
}

static int run_job(job *j)
{
    const int P = j->nparts;
    pthread_t *th = calloc((size_t)P, sizeof(pthread_t));
    char *started = calloc((size_t)P, 1);
    task *ts = calloc((size_t)P, sizeof(task));
    if (!th || !ts || !started) { free(th); free(ts); free(started); return -1; }
    /* chunk tasks: one thread per partition (a thread that cannot start runs inline) */
    for (int p = 0; p < P; ++p) {
        ts[p].j = j;
        ts[p].part = p;
        started[p] = P > 1 && pthread_create(&th[p], NULL, run_chunk, &ts[p]) == 0;
        if (!started[p]) run_chunk(&ts[p]);
    }
    for (int p = 0; p < P; ++p)
        if (started[p]) pthread_join(th[p], NULL);
    /* combine tree, split_every = 2: level by level, each merge one task */
    for (int stride = 1; stride < P; stride *= 2) {
        int k = 0;
        for (int p = 0; p + stride < P; p += 2 * stride, ++k) {
            ts[k].j = j;
            ts[k].part = p;
            ts[k].stride = stride;
            started[k] = pthread_create(&th[k], NULL, run_combine, &ts[k]) == 0;
            if (!started[k]) run_combine(&ts[k]);
        }
        for (int q = 0; q < k; ++q)
            if (started[q]) pthread_join(th[q], NULL);
    }
    free(th);
    free(ts);
    free(started);
    return 0;
}

static int alloc_parts(job *j, int want_v, int want_sum)
{
    j->count = calloc((size_t)j->nparts, sizeof(void *));
    j->countv = want_v ? calloc((size_t)j->nparts, sizeof(void *)) : NULL;
    j->sum = want_sum ? calloc((size_t)j->nparts, sizeof(void *)) : NULL;
    return j->count ? 0 : -1;
}

static void free_parts(job *j)
{
    for (int p = 1; p < j->nparts; ++p) {
        free(j->count[p]);
        if (j->countv) free(j->countv[p]);
        if (j->sum) free(j->sum[p]);
    }
    free(j->count);
    free(j->countv);
    free(j->sum);
}

/* Partition 0 aggregates into the caller's planes; the others into private ones. */
static int private_planes(job *j)
{
    for (int p = 1; p < j->nparts; ++p) {
        j->count[p] = malloc((size_t)j->ncells * 4);
        if (!j->count[p]) return -1;
        if (j->countv && !(j->countv[p] = malloc((size_t)j->ncells * 4))) return -1;
        if (j->sum && !(j->sum[p] = malloc((size_t)j->ncells * 8))) return -1;
    }
    return 0;
}

/* count() + mean(v) over float32 x, y, v: count, countv and the _sum_zero plane are
 * written; mean (sum / countv, NaN where empty) is finalized into `mean` when it is not
 * NULL. */
int cpub_points_count_mean_f32(const float *x, const float *y, const float *v, int64_t n,
                               const cpub_view *view, int32_t W, int32_t H, int32_t nthreads,
                               uint32_t *count, uint32_t *countv, double *sum, double *mean)
{
    job j;
    memset(&j, 0, sizeof j);
    j.kind = K_COUNT_MEAN; j.x = x; j.y = y; j.v = v; j.n = n; j.view = *view;
    j.W = W; j.H = H; j.ncat = 1; j.ncells = (int64_t)W * H;
    j.nparts = nthreads < 1 ? 1 : nthreads;
    if (alloc_parts(&j, 1, 1)) return -1;
    j.count[0] = count; j.countv[0] = countv; j.sum[0] = sum;
    int rc = private_planes(&j);
    if (!rc) rc = run_job(&j);
    if (!rc && mean)  /* finalize (reductions.py:1294-1297), once */
        for (int64_t c = 0; c < j.ncells; ++c)
            mean[c] = countv[c] > 0 ? sum[c] / (double)countv[c] : NAN;
    free_parts(&j);
    return rc;
}

int cpub_points_count_f64(const double *x, const double *y, int64_t n, const cpub_view *view,
                          int32_t W, int32_t H, int32_t nthreads, uint32_t *count)
{
    job j;
    memset(&j, 0, sizeof j);
    j.kind = K_COUNT_F64; j.x = x; j.y = y; j.n = n; j.view = *view;
    j.W = W; j.H = H; j.ncat = 1; j.ncells = (int64_t)W * H;
    j.nparts = nthreads < 1 ? 1 : nthreads;
    if (alloc_parts(&j, 0, 0)) return -1;
    j.count[0] = count;
    int rc = private_planes(&j);
    if (!rc) rc = run_job(&j);
    free_parts(&j);
    return rc;
}

/* by(cat, count()) with uint8 category codes (codes >= ncat drop the row) */
int cpub_points_count_cat_f32(const float *x, const float *y, const uint8_t *cat, int64_t n,
                              const cpub_view *view, int32_t W, int32_t H, int32_t ncat,
                              int32_t nthreads, uint32_t *count)
{
    job j;
    memset(&j, 0, sizeof j);
    j.kind = K_COUNT_CAT; j.x = x; j.y = y; j.v = cat; j.n = n; j.view = *view;
    j.W = W; j.H = H; j.ncat = ncat; j.ncells = (int64_t)W * H * ncat;
    j.nparts = nthreads < 1 ? 1 : nthreads;
    if (alloc_parts(&j, 0, 0)) return -1;
    j.count[0] = count;
    int rc = private_planes(&j);
    if (!rc) rc = run_job(&j);
    free_parts(&j);
    return rc;
}

/* LinesAxis1XConstant count: shared float64 x[ncols], float32 ys[nrows][ncols] */
int cpub_lines_xconst_f32(const double *xs, int64_t ncols, const float *ys, int64_t nrows,
                          const cpub_view *view, int32_t W, int32_t H, int32_t nthreads,
                          uint32_t *count)
{
    job j;
    memset(&j, 0, sizeof j);
    j.kind = K_LINES; j.x = xs; j.y = ys; j.n = nrows; j.ncols = ncols; j.view = *view;
    j.W = W; j.H = H; j.ncat = 1; j.ncells = (int64_t)W * H;
    j.nparts = nthreads < 1 ? 1 : nthreads;
    if (alloc_parts(&j, 0, 0)) return -1;
    j.count[0] = count;
    int rc = private_planes(&j);
    if (!rc) rc = run_job(&j);
    free_parts(&j);
    return rc;
}
