// dsx_lines.hip -- zero-width line rasterisation (glyphs/line.py, line_width == 0).
//
// One work item per segment.  Each item clips its segment to the view (Liang-Barsky,
// line.py:734-800), snaps both ends to pixels (map_onto_pixel_snap, 688-720) and walks
// the integer Bresenham path (986-1031), appending every pixel to the accumulator
// planes.  The arithmetic follows the reference's numba typing: coordinates promote to
// float64 except the segment delta x1 - x0, which a float32 column computes in float32.
#include "dsx_common.h"

namespace dsx {

struct LineArgs {
    int64_t nlines, nverts;
    const void *x, *y, *val;
    Cat cat;
    int64_t xls, xvs, yls, yvs;
    int32_t axis, plot_start;
    int32_t W, H, xlog, ylog;
    double sx, tx, sy, ty, xmin, xmax, ymin, ymax;
    int64_t xxmax, yymax;
};

// _clipt, line.py:783-800
__device__ __forceinline__ bool clipt(double p, double q, double &t0, double &t1) {
    if (p < 0 && q < 0) {
        const double r = q / p;
        if (r > t1) return false;
        if (r > t0) t0 = r;
    } else if (p > 0 && q < p) {
        const double r = q / p;
        if (r < t0) return false;
        if (r < t1) t1 = r;
    } else if (q < 0) {
        return false;
    }
    return true;
}

template <typename TV, bool HAS_VAL>
struct Sink {
    const Planes *p;
    int64_t W, H, ncat, c;
    TV v;
    __device__ __forceinline__ void operator()(int64_t x, int64_t y) const {
        if (x < 0) x += W;  // numba's negative-index wraparound
        if (y < 0) y += H;
        if ((uint64_t)x >= (uint64_t)W || (uint64_t)y >= (uint64_t)H) return;
        append<TV, HAS_VAL>(*p, (y * W + x) * ncat + c, v);
    }
};

// _bresenham, line.py:986-1031
template <typename S>
__device__ __forceinline__ void bresenham(const S &put, bool segment_start, int64_t x0,
                                          int64_t x1, int64_t y0, int64_t y1, bool clipped) {
    int64_t dx = x1 - x0;
    const int64_t ix = (dx > 0) - (dx < 0);
    dx = (dx < 0 ? -dx : dx) * 2;
    int64_t dy = y1 - y0;
    const int64_t iy = (dy > 0) - (dy < 0);
    dy = (dy < 0 ? -dy : dy) * 2;
    if (!clipped && !(dx | dy)) {
        put(x0, y0);
        return;
    }
    if (segment_start) put(x0, y0);
    if (dx >= dy) {
        int64_t err = 2 * dy - dx;
        while (x0 != x1) {
            if (err >= 0 && (err || ix > 0)) {
                err -= 2 * dx;
                y0 += iy;
            }
            err += 2 * dy;
            x0 += ix;
            put(x0, y0);
        }
    } else {
        int64_t err = 2 * dx - dy;
        while (y0 != y1) {
            if (err >= 0 && (err || iy > 0)) {
                err -= 2 * dy;
                x0 += ix;
            }
            err += 2 * dx;
            y0 += iy;
            put(x0, y0);
        }
    }
}

__device__ __forceinline__ int64_t snap1(double v, double s, double t, int64_t vmax,
                                         int32_t log) {
    const int64_t vv = (int64_t)__dadd_rn(__dmul_rn(axis_map(v, log), s), t);
    return vv == vmax ? vv - 1 : vv;
}

// draw_segment, line.py:1045-1095, for coordinates of column types TX / TY.
template <typename TX, typename TY, typename S>
__device__ __forceinline__ void draw_segment(const LineArgs &a, const S &put, bool segment_start,
                                             TX x0c, TX x1c, TY y0c, TY y1c) {
    double x0 = (double)x0c, x1 = (double)x1c, y0 = (double)y0c, y1 = (double)y1c;
    bool skip = (x0 != x0) || (y0 != y0) || (x1 != x1) || (y1 != y1);
    if (x0 < a.xmin && x1 < a.xmin) skip = true;
    else if (x0 > a.xmax && x1 > a.xmax) skip = true;
    else if (y0 < a.ymin && y1 < a.ymin) skip = true;
    else if (y0 > a.ymax && y1 > a.ymax) skip = true;
    double t0 = 0.0, t1 = 1.0;
    const double dx1 = (double)(TX)(x1c - x0c);  // column precision (numba typing)
    if (!clipt(-dx1, __dsub_rn(x0, a.xmin), t0, t1)) skip = true;
    if (!clipt(dx1, __dsub_rn(a.xmax, x0), t0, t1)) skip = true;
    const double dy1 = (double)(TY)(y1c - y0c);
    if (!clipt(-dy1, __dsub_rn(y0, a.ymin), t0, t1)) skip = true;
    if (!clipt(dy1, __dsub_rn(a.ymax, y0), t0, t1)) skip = true;
    if (skip) return;
    bool clipped = false;
    if (t1 < 1) {
        clipped = true;
        x1 = __dadd_rn(x0, __dmul_rn(t1, dx1));
        y1 = __dadd_rn(y0, __dmul_rn(t1, dy1));
    }
    if (t0 > 0) {
        clipped = true;
        segment_start = true;
        x0 = __dadd_rn(x0, __dmul_rn(t0, dx1));
        y0 = __dadd_rn(y0, __dmul_rn(t0, dy1));
    }
    const int64_t px0 = snap1(x0, a.sx, a.tx, a.xxmax, a.xlog);
    const int64_t py0 = snap1(y0, a.sy, a.ty, a.yymax, a.ylog);
    const int64_t px1 = snap1(x1, a.sx, a.tx, a.xxmax, a.xlog);
    const int64_t py1 = snap1(y1, a.sy, a.ty, a.yymax, a.ylog);
    // Clipped ends project into [0, W] x [0, H] (one pixel of rounding slack either
    // side); anything else comes from a degenerate view (infinite or NaN scale) and
    // would make the walk unbounded, so such a segment is not drawn.
    if (px0 < -1 || px0 > a.W + 1 || px1 < -1 || px1 > a.W + 1 || py0 < -1 ||
        py0 > a.H + 1 || py1 < -1 || py1 > a.H + 1)
        return;
    bresenham(put, segment_start, px0, px1, py0, py1, clipped);
}

template <typename TX, typename TY, typename TV, bool HAS_VAL>
__global__ __launch_bounds__(256) void lines_kernel(LineArgs a, Planes p) {
    const int64_t nseg = a.nverts - 1;
    const int64_t total = a.nlines * nseg;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const TX *X = reinterpret_cast<const TX *>(a.x);
    const TY *Y = reinterpret_cast<const TY *>(a.y);
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total; s += stride) {
        const int64_t l = s / nseg, k = s - l * nseg;
        const TX x0 = X[l * a.xls + k * a.xvs], x1 = X[l * a.xls + (k + 1) * a.xvs];
        const TY y0 = Y[l * a.yls + k * a.yvs], y1 = Y[l * a.yls + (k + 1) * a.yvs];
        bool start;
        if (k == 0) {
            start = a.axis == 0 ? (a.plot_start != 0) : true;
        } else {
            const TX xp = X[l * a.xls + (k - 1) * a.xvs];
            const TY yp = Y[l * a.yls + (k - 1) * a.yvs];
            start = (xp != xp) || (yp != yp);
        }
        const int64_t row = a.axis == 0 ? k : l;
        Sink<TV, HAS_VAL> put;
        put.p = &p; put.W = a.W; put.H = a.H; put.ncat = a.cat.ncat;
        put.v = TV(0);
        put.c = category(a.cat, row);
        if (put.c < 0) continue;
        if constexpr (HAS_VAL) put.v = reinterpret_cast<const TV *>(a.val)[row];
        draw_segment<TX, TY>(a, put, start, x0, x1, y0, y1);
    }
}

template <typename TX, typename TY>
static void launch_lines(const LineArgs &a, const Planes &p, int32_t vdt, int grid, int block,
                         hipStream_t s) {
    if (!a.val) lines_kernel<TX, TY, float, false><<<grid, block, 0, s>>>(a, p);
    else if (vdt == DSX_F32) lines_kernel<TX, TY, float, true><<<grid, block, 0, s>>>(a, p);
    else lines_kernel<TX, TY, double, true><<<grid, block, 0, s>>>(a, p);
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_lines_agg(const dsx_lines_desc *d, const dsx_view *v, const dsx_planes *acc,
                             void *stream) {
    if (!d || !v || !acc) return DSX_ERR_NULL;
    if (d->nlines < 0 || d->nverts < 0 || v->width <= 0 || v->height <= 0) return DSX_ERR_SHAPE;
    if (d->axis != 0 && d->axis != 1) return DSX_ERR_SHAPE;
    if ((d->x_dtype != DSX_F32 && d->x_dtype != DSX_F64) ||
        (d->y_dtype != DSX_F32 && d->y_dtype != DSX_F64))
        return DSX_ERR_DTYPE;
    if (d->val && d->val_dtype != DSX_F32 && d->val_dtype != DSX_F64) return DSX_ERR_DTYPE;
    const int cst = check_cat(d->cat);
    if (cst) return cst;
    const bool needs_val = acc->countv || acc->sum || acc->min || acc->max;
    if (needs_val && !d->val) return DSX_ERR_ACCUM;
    if (d->nlines == 0 || d->nverts < 2) return DSX_OK;
    if (!d->x || !d->y) return DSX_ERR_NULL;
    LineArgs a;
    a.nlines = d->nlines; a.nverts = d->nverts;
    a.x = d->x; a.y = d->y; a.val = d->val; a.cat = cat_of(d->cat);
    a.xls = d->x_line_stride; a.xvs = d->x_vert_stride;
    a.yls = d->y_line_stride; a.yvs = d->y_vert_stride;
    a.axis = d->axis; a.plot_start = d->plot_start;
    a.W = v->width; a.H = v->height; a.xlog = v->xlog; a.ylog = v->ylog;
    a.sx = v->sx; a.tx = v->tx; a.sy = v->sy; a.ty = v->ty;
    a.xmin = v->xmin; a.xmax = v->xmax; a.ymin = v->ymin; a.ymax = v->ymax;
    a.xxmax = v->xxmax; a.yymax = v->yymax;
    const Planes p = planes_of(acc);
    const int block = 256;
    const int grid = grid_for(d->nlines * (d->nverts - 1), block);
    hipStream_t s = as_stream(stream);
    const bool x32 = d->x_dtype == DSX_F32, y32 = d->y_dtype == DSX_F32;
    if (x32 && y32) launch_lines<float, float>(a, p, d->val_dtype, grid, block, s);
    else if (x32) launch_lines<float, double>(a, p, d->val_dtype, grid, block, s);
    else if (y32) launch_lines<double, float>(a, p, d->val_dtype, grid, block, s);
    else launch_lines<double, double>(a, p, d->val_dtype, grid, block, s);
    return launch_status();
}
