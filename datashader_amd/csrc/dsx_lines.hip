// dsx_lines.hip -- zero-width line rasterisation (glyphs/line.py, line_width == 0).
//
// General layouts: one work item per segment.  Each item clips its segment to the view (Liang-Barsky,
// line.py:734-800), snaps both ends to pixels (map_onto_pixel_snap, 688-720) and walks
// the integer Bresenham path (986-1031), appending every pixel to the accumulator
// planes.  The arithmetic follows the reference's numba typing: coordinates promote to
// float64 except the segment delta x1 - x0, which a float32 column computes in float32.
#include "dsx_common.h"

namespace dsx {

struct LineArgs {
    int64_t nlines, nverts;
    const void *x, *y, *val;
    const double *mean;  // squared-deviation mode (var / std second pass)
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

template <typename TV, bool HAS_VAL, bool SQDEV = false>
struct Sink {
    const Planes *p;
    const double *mean;
    int64_t W, H, ncat, c;
    TV v;
    __device__ __forceinline__ void operator()(int64_t x, int64_t y) const {
        if (x < 0) x += W;  // numba's negative-index wraparound
        if (y < 0) y += H;
        if ((uint64_t)x >= (uint64_t)W || (uint64_t)y >= (uint64_t)H) return;
        const int64_t cell = (y * W + x) * ncat + c;
        if constexpr (SQDEV) {
            if (v != v) return;
            const double d = (double)v - mean[cell];
            unsafeAtomicAdd(p->sum + cell, d * d);
        } else {
            append<TV, HAS_VAL>(*p, cell, v);
        }
    }
    template <typename I>
    __device__ __forceinline__ void walk(bool start, I x0, I x1, I y0, I y1, bool clipped) const;
};

// _bresenham, line.py:986-1031, in integer type I (the ends passed in are within one
// pixel of the canvas, so 32 bits hold every intermediate of the walk)
template <typename I, typename S>
__device__ __forceinline__ void bresenham(const S &put, bool segment_start, I x0, I x1, I y0,
                                          I y1, bool clipped) {
    I dx = x1 - x0;
    const I ix = (dx > 0) - (dx < 0);
    dx = (dx < 0 ? -dx : dx) * 2;
    I dy = y1 - y0;
    const I iy = (dy > 0) - (dy < 0);
    dy = (dy < 0 ? -dy : dy) * 2;
    if (!clipped && !(dx | dy)) {
        put(x0, y0);
        return;
    }
    if (segment_start) put(x0, y0);
    if (dx >= dy) {
        I err = 2 * dy - dx;
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
        I err = 2 * dx - dy;
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

template <typename TV, bool HAS_VAL, bool SQDEV>
template <typename I>
__device__ __forceinline__ void Sink<TV, HAS_VAL, SQDEV>::walk(bool start, I x0, I x1, I y0,
                                                              I y1, bool clipped) const {
    bresenham<I>(*this, start, x0, x1, y0, y1, clipped);
}

__device__ __forceinline__ int64_t snap1(double v, double s, double t, int64_t vmax,
                                         int32_t log) {
    const int64_t vv = (int64_t)__dadd_rn(__dmul_rn(axis_map(v, log), s), t);
    return vv == vmax ? vv - 1 : vv;
}

// draw_segment, line.py:1045-1095, for coordinates of column types TX / TY.
template <typename TX, typename TY, typename I, typename S>
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
    const double dy1 = (double)(TY)(y1c - y0c);
    // Both ends inside the view and each delta equal to the float64 difference of its
    // ends: then every _clipt call leaves (t0, t1) = (0, 1) and they are skipped.  (With
    // x1 <= xmax, the exact xmax - x0 is >= the exact x1 - x0, so the rounded q = xmax - x0
    // is >= p = fl64(x1 - x0); _clipt's q < p branch cannot fire, and likewise for the
    // other three bounds.)  A float32 delta that rounds away from the float64 one can
    // exceed q when a vertex lies on a bound; _clipt then clips (t1 < 1), so such a
    // segment takes the four calls.
    const bool inside = a.xmin <= x0 && x0 <= a.xmax && a.xmin <= x1 && x1 <= a.xmax &&
                        a.ymin <= y0 && y0 <= a.ymax && a.ymin <= y1 && y1 <= a.ymax &&
                        dx1 == __dsub_rn(x1, x0) && dy1 == __dsub_rn(y1, y0);
    if (!inside) {
        if (!clipt(-dx1, __dsub_rn(x0, a.xmin), t0, t1)) skip = true;
        if (!clipt(dx1, __dsub_rn(a.xmax, x0), t0, t1)) skip = true;
        if (!clipt(-dy1, __dsub_rn(y0, a.ymin), t0, t1)) skip = true;
        if (!clipt(dy1, __dsub_rn(a.ymax, y0), t0, t1)) skip = true;
    }
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
    put.template walk<I>(segment_start, (I)px0, (I)px1, (I)py0, (I)py1, clipped);
}

template <typename TX, typename TY, typename TV, bool HAS_VAL, bool SQDEV>
__global__ __launch_bounds__(256) void lines_kernel(LineArgs a, Planes p) {
    const int64_t nseg = a.nverts - 1;
    const int64_t total = a.nlines * nseg;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const TX *X = reinterpret_cast<const TX *>(a.x);
    const TY *Y = reinterpret_cast<const TY *>(a.y);
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total; s += stride) {
        const int64_t l = s / nseg, k = s - l * nseg;
        if (k == 0 && a.axis == 0 && a.plot_start == DSX_PLOT_CONTEXT) continue;  // context vertex
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
        Sink<TV, HAS_VAL, SQDEV> put;
        put.p = &p; put.mean = a.mean; put.W = a.W; put.H = a.H; put.ncat = a.cat.ncat;
        put.v = TV(0);
        put.c = category(a.cat, row);
        if (put.c < 0) continue;
        if constexpr (HAS_VAL) put.v = reinterpret_cast<const TV *>(a.val)[row];
        draw_segment<TX, TY, int64_t>(a, put, start, x0, x1, y0, y1);
    }
}

// ---- x-constant series, count only: LDS strips ---------------------------------------
// Lines of an axis-1 frame whose x is shared (x_line_stride == 0, LinesAxis1XConstant,
// line.py:334-380) put segment j of every line between the same two x values, so the
// pixels of segments [j0, j0 + k) of all lines fall in one narrow band of columns.
// Workgroup (s, b) rasterises segments [s*k, s*k + k) of lines [b*L, b*L + L) into an
// LDS copy of that band (all H rows, row pitch odd so that rows spread over the LDS
// banks), then adds the band's non-zero cells to the count plane with one global atomic
// per cell.  A pixel outside the band (x not monotone, wraparound) goes straight to the
// global plane, so the result never depends on the band guess.
constexpr int kStripThreads = 1024;
constexpr int kStripLds = 160 * 1024;
constexpr int kStripMaxK = 8;        // segments per strip

// The band, [H][pitch] counters in the workgroup's dynamic LDS.  Named at file scope so
// that every access through the sink below compiles to an LDS (ds_*) atomic.
extern __shared__ uint32_t strip_band[];

struct StripSink {
    uint32_t *count;
    int32_t W, H, c0, cols, pitch;
    __device__ __forceinline__ void operator()(int32_t x, int32_t y) const {
        if (x < 0) x += W;  // numba's negative-index wraparound
        if (y < 0) y += H;
        if ((uint32_t)x >= (uint32_t)W || (uint32_t)y >= (uint32_t)H) return;
        const uint32_t cx = (uint32_t)(x - c0);
        if (cx < (uint32_t)cols) atomicAdd(&strip_band[y * pitch + cx], 1u);
        else atomicAdd(count + (int64_t)y * W + x, 1u);
    }
    // The pixels of a walk lie in the box of its ends.  When that box is inside the band
    // and the canvas, the walk runs without per-pixel checks as one loop over the major
    // axis (the two loops of _bresenham with the axes named by role: same pixels, same
    // tie rule); otherwise it is the checked walk.
    template <typename I>
    __device__ __forceinline__ void walk(bool start, I x0, I x1, I y0, I y1, bool clipped) const {
        const I xl = x0 < x1 ? x0 : x1, xh = x0 < x1 ? x1 : x0;
        const I yl = y0 < y1 ? y0 : y1, yh = y0 < y1 ? y1 : y0;
        if (!(xl >= c0 && xh < c0 + cols && xh < W && yl >= 0 && yh < H)) {
            bresenham<I>(*this, start, x0, x1, y0, y1, clipped);
            return;
        }
        uint32_t *base = strip_band - c0;  // band cell of pixel (x, y): base[y * pitch + x]
        I dx = x1 - x0;
        const I ix = (dx > 0) - (dx < 0);
        dx = (dx < 0 ? -dx : dx) * 2;
        I dy = y1 - y0;
        const I iy = (dy > 0) - (dy < 0);
        dy = (dy < 0 ? -dy : dy) * 2;
        if (!clipped && !(dx | dy)) {
            atomicAdd(&base[y0 * pitch + x0], 1u);
            return;
        }
        if (start) atomicAdd(&base[y0 * pitch + x0], 1u);
        const bool xmajor = dx >= dy;
        I a = xmajor ? x0 : y0, a1 = xmajor ? x1 : y1, b = xmajor ? y0 : x0;
        const I ia = xmajor ? ix : iy, ib = xmajor ? iy : ix;
        const I da = xmajor ? dx : dy, db = xmajor ? dy : dx;
        // LDS offsets of one step along each axis
        const I sa = xmajor ? ia : ia * pitch, sb = xmajor ? ib * pitch : ib;
        I off = y0 * pitch + x0;
        I err = 2 * db - da;
        while (a != a1) {
            if (err >= 0 && (err || ia > 0)) {
                err -= 2 * da;
                b += ib;
                off += sb;
            }
            err += 2 * db;
            a += ia;
            off += sa;
            atomicAdd(&base[off], 1u);
        }
        (void)b;
    }
};

template <typename TX, typename TY>
__global__ __launch_bounds__(kStripThreads) void lines_xconst_strip_kernel(LineArgs a,
                                                                           uint32_t *count,
                                                                           int32_t k,
                                                                           int32_t cols,
                                                                           int32_t pitch,
                                                                           int64_t lpb) {
    uint32_t *band = strip_band;
    __shared__ int32_t s_c0;
    const TX *X = reinterpret_cast<const TX *>(a.x);
    const TY *Y = reinterpret_cast<const TY *>(a.y);
    const int64_t nseg = a.nverts - 1;
    const int64_t j0 = (int64_t)blockIdx.x * k;
    const int64_t j1 = j0 + k < nseg ? j0 + k : nseg;  // segments [j0, j1)
    const int64_t l0 = (int64_t)blockIdx.y * lpb;
    const int64_t l1 = l0 + lpb < a.nlines ? l0 + lpb : a.nlines;
    if (threadIdx.x == 0) {
        // first column of the band: the leftmost snapped vertex of the strip, with x
        // clamped to the view as clipping would (a clipped end lies between its vertices)
        int64_t lo = a.W;
        for (int64_t j = j0; j <= j1; ++j) {
            double x = (double)X[j * a.xvs];
            if (x != x) continue;
            x = x < a.xmin ? a.xmin : (x > a.xmax ? a.xmax : x);
            const int64_t c = snap1(x, a.sx, a.tx, a.xxmax, a.xlog);
            lo = c < lo ? c : lo;
        }
        s_c0 = (int32_t)(lo < 0 ? 0 : (lo > a.W ? a.W : lo));
    }
    for (int i = threadIdx.x; i < a.H * pitch; i += kStripThreads) band[i] = 0u;
    __syncthreads();
    StripSink put;
    put.count = count; put.W = a.W; put.H = a.H;
    put.c0 = s_c0; put.cols = cols; put.pitch = pitch;
    const int nv = (int)(j1 - j0) + 1;  // vertices j0 .. j1
    TX xs[kStripMaxK + 1];
#pragma unroll
    for (int q = 0; q <= kStripMaxK; ++q) xs[q] = q < nv ? X[(j0 + q) * a.xvs] : TX(0);
    const bool xprev_nan = j0 > 0 && X[(j0 - 1) * a.xvs] != X[(j0 - 1) * a.xvs];
    for (int64_t l = l0 + threadIdx.x; l < l1; l += kStripThreads) {
        const TY *Yl = Y + l * a.yls;
        TY ys[kStripMaxK + 1];
#pragma unroll
        for (int q = 0; q <= kStripMaxK; ++q) ys[q] = q < nv ? Yl[(j0 + q) * a.yvs] : TY(0);
        // segment_start (line.py:1371-1373): the first segment of a line, or one whose
        // previous vertex is missing
        bool start = true;
        if (j0 > 0) {
            const TY yp = Yl[(j0 - 1) * a.yvs];
            start = xprev_nan || (yp != yp);
        }
        // segments drawn from the front of shifting register windows (a rolled loop with
        // constant register indices: one copy of the rasteriser, no indexed registers)
        TX xq[kStripMaxK + 1];
#pragma unroll
        for (int q = 0; q <= kStripMaxK; ++q) xq[q] = xs[q];
#pragma unroll 1
        for (int q = 0; q + 1 < nv; ++q) {
            draw_segment<TX, TY, int32_t>(a, put, start, xq[0], xq[1], ys[0], ys[1]);
            start = (xq[0] != xq[0]) || (ys[0] != ys[0]);
#pragma unroll
            for (int i = 0; i < kStripMaxK; ++i) {
                xq[i] = xq[i + 1];
                ys[i] = ys[i + 1];
            }
        }
    }
    __syncthreads();
    const int32_t c0 = s_c0;
    for (int i = threadIdx.x; i < a.H * pitch; i += kStripThreads) {
        const uint32_t v = band[i];
        if (!v) continue;
        const int32_t y = i / pitch, cx = i - y * pitch;
        if (cx < cols && c0 + cx < a.W) atomicAdd(count + (int64_t)y * a.W + c0 + cx, v);
    }
}

template <typename TX, typename TY>
static void launch_lines(const LineArgs &a, const Planes &p, int32_t vdt, int grid, int block,
                         hipStream_t s) {
    if (!a.val) lines_kernel<TX, TY, float, false, false><<<grid, block, 0, s>>>(a, p);
    else if (a.mean && vdt == DSX_F32) lines_kernel<TX, TY, float, true, true><<<grid, block, 0, s>>>(a, p);
    else if (a.mean) lines_kernel<TX, TY, double, true, true><<<grid, block, 0, s>>>(a, p);
    else if (vdt == DSX_F32) lines_kernel<TX, TY, float, true, false><<<grid, block, 0, s>>>(a, p);
    else lines_kernel<TX, TY, double, true, false><<<grid, block, 0, s>>>(a, p);
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
    if (acc->mean) return DSX_ERR_ACCUM;  // lines accumulate: a mean is finalized afterwards
    const bool needs_val = acc->countv || acc->sum || acc->min || acc->max;
    if (needs_val && !d->val) return DSX_ERR_ACCUM;
    if (d->mean && (!d->val || !acc->sum || acc->count || acc->countv || acc->min || acc->max))
        return DSX_ERR_ACCUM;
    if (d->nlines == 0 || d->nverts < 2) return DSX_OK;
    if (!d->x || !d->y) return DSX_ERR_NULL;
    LineArgs a;
    a.nlines = d->nlines; a.nverts = d->nverts;
    a.x = d->x; a.y = d->y; a.val = d->val; a.cat = cat_of(d->cat); a.mean = d->mean;
    a.xls = d->x_line_stride; a.xvs = d->x_vert_stride;
    a.yls = d->y_line_stride; a.yvs = d->y_vert_stride;
    a.axis = d->axis; a.plot_start = d->plot_start;
    a.W = v->width; a.H = v->height; a.xlog = v->xlog; a.ylog = v->ylog;
    a.sx = v->sx; a.tx = v->tx; a.sy = v->sy; a.ty = v->ty;
    a.xmin = v->xmin; a.xmax = v->xmax; a.ymin = v->ymin; a.ymax = v->ymax;
    a.xxmax = v->xxmax; a.yymax = v->yymax;
    const Planes p = planes_of(acc);
    hipStream_t s = as_stream(stream);
    const bool x32 = d->x_dtype == DSX_F32, y32 = d->y_dtype == DSX_F32;
    // x-constant series with count / any only: LDS strips
    int64_t pitch = kStripLds / (4 * (int64_t)v->height);
    if (pitch % 2 == 0) --pitch;
    if (d->axis == 1 && d->x_line_stride == 0 && !d->val && d->cat.kind == DSX_CAT_NONE &&
        acc->count && !needs_val && pitch >= 5) {
        const int64_t nseg = d->nverts - 1;
        int64_t k = (pitch - 2) * nseg / v->width;  // segments whose band fits, for x spread evenly
        k = k < 1 ? 1 : (k > kStripMaxK ? kStripMaxK : k);
        const int64_t nstrips = (nseg + k - 1) / k;
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) {
            int c = 0;
            if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
                c > 0)
                cus = c;
        }
        int64_t nblocks = (cus + nstrips - 1) / nstrips;
        if (nblocks > d->nlines) nblocks = d->nlines;
        if (nblocks < 1) nblocks = 1;
        if (nstrips <= 65535 && nblocks <= 65535) {
            const int64_t lpb = (d->nlines + nblocks - 1) / nblocks;
            const size_t lds = (size_t)4 * v->height * pitch;
            auto kern = lines_xconst_strip_kernel<float, float>;
            if (!x32 && y32) kern = lines_xconst_strip_kernel<double, float>;
            else if (x32 && !y32) kern = lines_xconst_strip_kernel<float, double>;
            else if (!x32 && !y32) kern = lines_xconst_strip_kernel<double, double>;
            const hipError_t e = hipFuncSetAttribute((const void *)kern,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     (int)lds);
            if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
            hipLaunchKernelGGL(kern, dim3((unsigned)nstrips, (unsigned)nblocks), dim3(kStripThreads),
                               lds, s, a, p.count, (int32_t)k, (int32_t)pitch, (int32_t)pitch, lpb);
            return launch_status();
        }
    }
    const int block = 256;
    const int grid = grid_for(d->nlines * (d->nverts - 1), block);
    if (x32 && y32) launch_lines<float, float>(a, p, d->val_dtype, grid, block, s);
    else if (x32) launch_lines<float, double>(a, p, d->val_dtype, grid, block, s);
    else if (y32) launch_lines<double, float>(a, p, d->val_dtype, grid, block, s);
    else launch_lines<double, double>(a, p, d->val_dtype, grid, block, s);
    return launch_status();
}
