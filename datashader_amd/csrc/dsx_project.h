// dsx_project.h -- the points projection shared by the points kernels (gfx950).
#pragma once

#include "dsx_common.h"

namespace dsx {

struct PointArgs {
    int64_t n;
    const void *x, *y, *val;
    Cat cat;
    int32_t W, H, xlog, ylog;
    double sx, tx, sy, ty, xmin, xmax, ymin, ymax;
};

// A value column is present when its pointer is set; an empty input (whose column has
// no storage) is identified by its dtype alone.
static bool has_value(const dsx_points_desc *d) {
    return d->val != nullptr ||
           (d->n == 0 && (d->val_dtype == DSX_F32 || d->val_dtype == DSX_F64));
}

// ---- projection -----------------------------------------------------------------------
// Pixel of a row from its coordinates (glyphs/points.py:197-206); false when dropped.
template <typename TC>
__device__ __forceinline__ bool locate_xy(const PointArgs &a, TC xc, TC yc, int32_t &px,
                                          int32_t &py) {
    const double x = (double)xc, y = (double)yc;
    if (!((a.xmin <= x && x <= a.xmax) && (a.ymin <= y && y <= a.ymax))) return false;
    int32_t xx = proj(axis_map(x, a.xlog), a.sx, a.tx);
    int32_t yy = proj(axis_map(y, a.ylog), a.sy, a.ty);
    xx = xx >= a.W ? a.W - 1 : xx;
    yy = yy >= a.H ? a.H - 1 : yy;
    if (xx < 0) xx += a.W;  // numba's negative-index wraparound
    if (yy < 0) yy += a.H;
    if ((uint32_t)xx >= (uint32_t)a.W || (uint32_t)yy >= (uint32_t)a.H) return false;
    px = xx; py = yy;
    return true;
}

// Pixel and category of row i; false when the row is dropped.
template <typename TC>
__device__ __forceinline__ bool locate(const PointArgs &a, int64_t i, TC xc, TC yc,
                                       int32_t &px, int32_t &py, int32_t &pc) {
    if (!locate_xy<TC>(a, xc, yc, px, py)) return false;
    const int64_t c = category(a.cat, i);
    if (c < 0) return false;
    pc = (int32_t)c;
    return true;
}

static int check_points(const dsx_points_desc *d, const dsx_view *v) {
    if (!d || !v) return DSX_ERR_NULL;
    if (d->n < 0 || v->width <= 0 || v->height <= 0) return DSX_ERR_SHAPE;
    if (v->width >= (1 << 24) || v->height >= (1 << 24)) return DSX_ERR_SHAPE;  // 24-bit math
    if (d->n >= (1ll << 32)) return DSX_ERR_SHAPE;  // spill positions are 32-bit per tile
    if (d->n > 0 && (!d->x || !d->y)) return DSX_ERR_NULL;
    if (d->xy_dtype != DSX_F32 && d->xy_dtype != DSX_F64) return DSX_ERR_DTYPE;
    if (d->val && d->val_dtype != DSX_F32 && d->val_dtype != DSX_F64) return DSX_ERR_DTYPE;
    if (d->strategy < DSX_STRATEGY_AUTO || d->strategy > DSX_STRATEGY_TILES) return DSX_ERR_MODE;
    return check_cat(d->cat);
}

// The n-pass form of the n-value reductions (defined with the row-index window pass in
// dsx_points.hip, run by dsx_points_nselect).
__attribute__((visibility("hidden"))) size_t ri_npass_workspace_size(const dsx_points_desc *d,
                                                                     const dsx_view *v);
__attribute__((visibility("hidden"))) int ri_npass(const dsx_points_desc *d, const dsx_view *v,
                                                   int mode, int nsel, int nan_check,
                                                   int64_t *index, void *workspace,
                                                   size_t workspace_bytes, hipStream_t st);

static PointArgs args_of(const dsx_points_desc *d, const dsx_view *v) {
    PointArgs a;
    a.n = d->n; a.x = d->x; a.y = d->y; a.val = d->val; a.cat = cat_of(d->cat);
    a.W = v->width; a.H = v->height; a.xlog = v->xlog; a.ylog = v->ylog;
    a.sx = v->sx; a.tx = v->tx; a.sy = v->sy; a.ty = v->ty;
    a.xmin = v->xmin; a.xmax = v->xmax; a.ymin = v->ymin; a.ymax = v->ymax;
    return a;
}

}  // namespace dsx
