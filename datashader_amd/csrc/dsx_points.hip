// dsx_points.hip -- point projection + per-bin aggregation (glyphs/points.py:188-221).
//
// One pass streams the x / y (/ value / category) columns once.  Each row is tested
// against the inclusive bounds, projected with the reference's exact float64
// arithmetic, and every requested accumulator plane is updated with a device-scope
// atomic at its bin.
#include "dsx_common.h"

namespace dsx {

struct PointArgs {
    int64_t n;
    const void *x, *y, *val;
    Cat cat;
    int32_t W, H, xlog, ylog;
    double sx, tx, sy, ty, xmin, xmax, ymin, ymax;
};

// Bin (cell index) of row i, or -1 when the row is dropped.
template <typename TC>
__device__ __forceinline__ int64_t point_cell(const PointArgs &a, int64_t i) {
    const double x = (double)reinterpret_cast<const TC *>(a.x)[i];
    const double y = (double)reinterpret_cast<const TC *>(a.y)[i];
    if (!((a.xmin <= x && x <= a.xmax) && (a.ymin <= y && y <= a.ymax))) return -1;
    int32_t xx = proj(axis_map(x, a.xlog), a.sx, a.tx);
    int32_t yy = proj(axis_map(y, a.ylog), a.sy, a.ty);
    xx = xx >= a.W ? a.W - 1 : xx;
    yy = yy >= a.H ? a.H - 1 : yy;
    if (xx < 0) xx += a.W;  // numba's negative-index wraparound
    if (yy < 0) yy += a.H;
    if ((uint32_t)xx >= (uint32_t)a.W || (uint32_t)yy >= (uint32_t)a.H) return -1;
    const int64_t c = category(a.cat, i);
    if (c < 0) return -1;
    return ((int64_t)yy * a.W + xx) * a.cat.ncat + c;
}

template <typename TC, typename TV, bool HAS_VAL>
__global__ __launch_bounds__(256) void points_agg_kernel(PointArgs a, Planes p) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        const int64_t cell = point_cell<TC>(a, i);
        if (cell < 0) continue;
        TV v = TV(0);
        if constexpr (HAS_VAL) v = reinterpret_cast<const TV *>(a.val)[i];
        append<TV, HAS_VAL>(p, cell, v);
    }
}

// Squared deviations from the per-bin mean, for var / std.
template <typename TC, typename TV>
__global__ __launch_bounds__(256) void points_m2_kernel(PointArgs a, const double *mean,
                                                        float *hi, float *lo, double *m64) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        const int64_t cell = point_cell<TC>(a, i);
        if (cell < 0) continue;
        const TV v = reinterpret_cast<const TV *>(a.val)[i];
        if (v != v) continue;
        const double d = (double)v - mean[cell];
        if constexpr (sizeof(TV) == 4) {
            const float df = (float)d;
            add_comp(hi + cell, lo + cell, df * df);
        } else {
            unsafeAtomicAdd(m64 + cell, d * d);
        }
    }
}

static int check_points(const dsx_points_desc *d, const dsx_view *v) {
    if (!d || !v) return DSX_ERR_NULL;
    if (d->n < 0 || v->width <= 0 || v->height <= 0) return DSX_ERR_SHAPE;
    if (d->n > 0 && (!d->x || !d->y)) return DSX_ERR_NULL;
    if (d->xy_dtype != DSX_F32 && d->xy_dtype != DSX_F64) return DSX_ERR_DTYPE;
    if (d->val && d->val_dtype != DSX_F32 && d->val_dtype != DSX_F64) return DSX_ERR_DTYPE;
    return check_cat(d->cat);
}

static PointArgs args_of(const dsx_points_desc *d, const dsx_view *v) {
    PointArgs a;
    a.n = d->n; a.x = d->x; a.y = d->y; a.val = d->val; a.cat = cat_of(d->cat);
    a.W = v->width; a.H = v->height; a.xlog = v->xlog; a.ylog = v->ylog;
    a.sx = v->sx; a.tx = v->tx; a.sy = v->sy; a.ty = v->ty;
    a.xmin = v->xmin; a.xmax = v->xmax; a.ymin = v->ymin; a.ymax = v->ymax;
    return a;
}

}  // namespace dsx

using namespace dsx;

extern "C" int dsx_points_agg(const dsx_points_desc *d, const dsx_view *v,
                              const dsx_accums *acc, void *stream) {
    int st = check_points(d, v);
    if (st) return st;
    if (!acc) return DSX_ERR_NULL;
    const bool f32v = d->val && d->val_dtype == DSX_F32;
    const bool needs_val = acc->countv || acc->sum_hi || acc->sum64 || acc->min32 ||
                           acc->max32 || acc->min64 || acc->max64;
    if (needs_val && !d->val) return DSX_ERR_ACCUM;
    if ((acc->sum_hi != nullptr) != (acc->sum_lo != nullptr)) return DSX_ERR_ACCUM;
    if (d->val && f32v && (acc->sum64 || acc->min64 || acc->max64)) return DSX_ERR_ACCUM;
    if (d->val && !f32v && (acc->sum_hi || acc->min32 || acc->max32)) return DSX_ERR_ACCUM;
    if (d->n == 0) return DSX_OK;
    const PointArgs a = args_of(d, v);
    const Planes p = planes_of(acc);
    const int block = 256;
    const int grid = grid_for(d->n, block);
    hipStream_t s = as_stream(stream);
    const bool c32 = d->xy_dtype == DSX_F32;
    if (!d->val) {
        if (c32) points_agg_kernel<float, float, false><<<grid, block, 0, s>>>(a, p);
        else points_agg_kernel<double, float, false><<<grid, block, 0, s>>>(a, p);
    } else if (f32v) {
        if (c32) points_agg_kernel<float, float, true><<<grid, block, 0, s>>>(a, p);
        else points_agg_kernel<double, float, true><<<grid, block, 0, s>>>(a, p);
    } else {
        if (c32) points_agg_kernel<float, double, true><<<grid, block, 0, s>>>(a, p);
        else points_agg_kernel<double, double, true><<<grid, block, 0, s>>>(a, p);
    }
    return launch_status();
}

extern "C" int dsx_points_m2(const dsx_points_desc *d, const dsx_view *v, const double *mean,
                             float *m2_hi, float *m2_lo, double *m2_64, void *stream) {
    int st = check_points(d, v);
    if (st) return st;
    if (!d->val || !mean) return DSX_ERR_NULL;
    const bool f32v = d->val_dtype == DSX_F32;
    if (f32v && (!m2_hi || !m2_lo)) return DSX_ERR_ACCUM;
    if (!f32v && !m2_64) return DSX_ERR_ACCUM;
    if (d->n == 0) return DSX_OK;
    const PointArgs a = args_of(d, v);
    const int block = 256;
    const int grid = grid_for(d->n, block);
    hipStream_t s = as_stream(stream);
    const bool c32 = d->xy_dtype == DSX_F32;
    if (f32v) {
        if (c32) points_m2_kernel<float, float><<<grid, block, 0, s>>>(a, mean, m2_hi, m2_lo, m2_64);
        else points_m2_kernel<double, float><<<grid, block, 0, s>>>(a, mean, m2_hi, m2_lo, m2_64);
    } else {
        if (c32) points_m2_kernel<float, double><<<grid, block, 0, s>>>(a, mean, m2_hi, m2_lo, m2_64);
        else points_m2_kernel<double, double><<<grid, block, 0, s>>>(a, mean, m2_hi, m2_lo, m2_64);
    }
    return launch_status();
}
