// dsx_shade.hip -- tf.shade on the device: aggregate planes -> RGBA pixels.
//
// The reference colours an aggregate on the host with numpy (transfer_functions/
// __init__.py:148-747): mask, offset, optional clip, a transform (linear, log1p, cube
// root or histogram equalisation), the span of the transformed values, and np.interp
// onto a colormap.  Here every step is a kernel over the plane and every scalar the
// reference derives from the data lives in a device-resident Params block that later
// kernels read, so one call is a fixed launch sequence with no host round trip.
//
// The arithmetic follows numpy's: values are offset in the plane's own dtype (float32
// stays float32), transforms of float32 planes round to float32, the histogram bins of
// np.histogram (uniform edges from np.linspace in the bin dtype, numpy's index
// correction at the edges), np.interp's search and interpolation formula, and the
// truncating float -> uint8 casts.
#include <type_traits>

#include "dsx_common.h"

namespace dsx {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxParts = 1024;   // blocks of a reduction pass (one partial each)
constexpr int kPlanThreads = 256;   // the single workgroup of a plan kernel
constexpr int kBins = 65536;      // eq_hist's nbins default (transfer_functions:148)
constexpr int kHalfBins = 32768;  // bins one histogram workgroup holds in LDS (128 KB)
constexpr int kCdfThreads = 1024;             // bins per cdf workgroup
constexpr int kCdfBlocks = kBins / kCdfThreads;  // 64
constexpr int kHistCells = 8192;  // cells per histogram workgroup
constexpr int kMaxKnots = 4096;   // colours of a list colormap / alpha steps

enum Kind : int32_t { K_INT = 0, K_F32 = 1, K_F64 = 2 };

struct Part {
    unsigned long long n;  // valid values
    double mn, mx;         // their min / max
    double pmin;           // smallest value > 0 (totals of unsigned planes)
    uint32_t zero;         // a valid value == 0 was seen
    uint32_t pad;
};

// Scalars derived on the device, written by the plan kernels, read by the passes.
struct Params {
    // raw plane (2D) or per-pixel totals (3D)
    unsigned long long n;
    double mn, mx, pmin;
    uint32_t zero;
    int32_t all_masked;
    double offset;          // value subtracted, in the working arithmetic
    int32_t kind;           // Kind of the offset values
    int32_t zero_masked;    // 3D: totals == 0 are masked (unsigned planes)
    double clip_lo, clip_hi;          // span bounds as given (compared in float64)
    double clip_lo_set, clip_hi_set;  // the bounds cast to the clipped array's dtype
    // offset values
    unsigned long long vn;
    double vmin, vmax;
    // eq_hist
    int32_t int_path, nb, L, levels;
    double first, last, denom, step, delta;
    int32_t step_zero, pad0;
    double total;
    // transformed values and the colour knots
    double tmin, tmax;
    int32_t kf32, nk, kmode, pad1;  // kmode 0: i*step + start, 1: i/div*delta + start, 2: n == 1
    double kstart, kstop, kstep, kdelta;
};

struct Ws {
    Params *P;
    Part *parts;
    uint32_t *hist;
    double *xp, *cdf;
    int32_t *rank;                   // position of a non-empty bin among the non-empty bins
    double *eslope;                  // np.interp slopes between consecutive cdf points
    double *kx, *ks;                 // colour knots and the slopes between them (3 / knot)
    unsigned long long *chunk_sum;   // per 1024-bin chunk: values
    uint32_t *chunk_nz;              // per 1024-bin chunk: non-empty bins
    double *val;  // H*W working plane (offset values, then transformed values)
};

__host__ __device__ inline size_t al256(size_t x) { return (x + 255) / 256 * 256; }

struct Layout {
    size_t o_params, o_parts, o_hist, o_xp, o_cdf, o_rank, o_chunk, o_eslope, o_kx, o_ks, o_val,
        total;
};

inline Layout layout_of(int64_t pixels) {
    Layout L;
    size_t o = 0;
    L.o_params = o; o = al256(o + sizeof(Params));
    L.o_parts = o; o = al256(o + sizeof(Part) * kMaxParts);
    L.o_hist = o; o = al256(o + 4 * (size_t)kBins);
    L.o_xp = o; o = al256(o + 8 * (size_t)kBins);
    L.o_cdf = o; o = al256(o + 8 * (size_t)kBins);
    L.o_rank = o; o = al256(o + 4 * (size_t)kBins);
    L.o_chunk = o; o = al256(o + 12 * (size_t)kCdfBlocks);
    L.o_eslope = o; o = al256(o + 8 * (size_t)kBins);
    L.o_kx = o; o = al256(o + 8 * (size_t)kMaxKnots);
    L.o_ks = o; o = al256(o + 24 * (size_t)kMaxKnots);
    L.o_val = o; o = al256(o + 8 * (size_t)pixels);
    L.total = o;
    return L;
}

// ---- dtype helpers -------------------------------------------------------------------

template <typename T> struct KindOf { static constexpr int value = K_INT; };
template <> struct KindOf<float> { static constexpr int value = K_F32; };
template <> struct KindOf<double> { static constexpr int value = K_F64; };

// _interpolate's mask (transfer_functions:268-275): bool -> ~data, unsigned -> data == 0,
// float -> isnan; signed integers have no missing value.
template <typename T, bool BOOLEAN>
__device__ __forceinline__ bool masked_raw(T v) {
    if constexpr (BOOLEAN) return v == 0;
    else if constexpr (std::is_same<T, uint32_t>::value) return v == 0u;
    else if constexpr (std::is_floating_point<T>::value) return v != v;
    else return false;
}

// A float64 value cast to an array dtype the way numpy / numba assign it: truncation to
// integers, round-to-nearest to float32.
__device__ __forceinline__ double cast_to(double x, int kind) {
    if (kind == K_INT) return trunc(x);
    if (kind == K_F32) return (double)(float)x;
    return x;
}

// float -> uint8 as numpy's astype(np.uint8) does on x86-64: convert to int32 with
// truncation (out of range and NaN give 0x80000000) and keep the low byte.
__device__ __forceinline__ uint32_t u8_of(double x) {
    if (!(x > -2147483649.0 && x < 2147483648.0)) return 0u;
    return (uint32_t)(int32_t)x & 0xFFu;
}

// ---- reductions -------------------------------------------------------------------------

__device__ __forceinline__ void part_init(Part &p) {
    p.n = 0; p.mn = __builtin_inf(); p.mx = -__builtin_inf(); p.pmin = __builtin_inf();
    p.zero = 0; p.pad = 0;
}

__device__ __forceinline__ void part_add(Part &p, double v) {
    p.n += 1;
    p.mn = fmin(p.mn, v);
    p.mx = fmax(p.mx, v);
    if (v > 0.0) p.pmin = fmin(p.pmin, v);
    if (v == 0.0) p.zero = 1;
}

__device__ __forceinline__ void part_merge(Part &a, const Part &b) {
    a.n += b.n;
    a.mn = fmin(a.mn, b.mn);
    a.mx = fmax(a.mx, b.mx);
    a.pmin = fmin(a.pmin, b.pmin);
    a.zero |= b.zero;
}

// Block-wide merge of per-thread partials; thread 0 writes the block's partial.
__device__ void block_part(Part p, Part *out) {
    __shared__ Part sp[kThreads];
    sp[threadIdx.x] = p;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part_merge(sp[threadIdx.x], sp[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sp[0];
}

// All partials of the previous pass merged (one block, every thread gets the result).
__device__ Part merge_parts(const Part *parts, int nparts) {
    __shared__ Part sp[kPlanThreads];
    Part p;
    part_init(p);
    for (int i = threadIdx.x; i < nparts; i += kPlanThreads) part_merge(p, parts[i]);
    sp[threadIdx.x] = p;
    __syncthreads();
    for (int s = kPlanThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part_merge(sp[threadIdx.x], sp[threadIdx.x + s]);
        __syncthreads();
    }
    Part r = sp[0];
    __syncthreads();
    return r;
}

// ---- np.linspace / np.interp ----------------------------------------------------------

// Knot i of np.linspace(start, stop, n) evaluated in float32 or float64
// (numpy/_core/function_base.py linspace: arange * step + start, last = stop).
__device__ __forceinline__ double knot(const Params &P, int i) {
    if (P.nk > 1 && i == P.nk - 1) return P.kstop;
    if (P.kf32) {
        float y = (float)i;
        if (P.kmode == 0) y = y * (float)P.kstep;
        else if (P.kmode == 1) { y = y / (float)(P.nk - 1); y = y * (float)P.kdelta; }
        else y = y * (float)P.kdelta;
        return (double)(y + (float)P.kstart);
    }
    double y = (double)i;
    if (P.kmode == 0) y = y * P.kstep;
    else if (P.kmode == 1) { y = y / (double)(P.nk - 1); y = y * P.kdelta; }
    else y = y * P.kdelta;
    return y + P.kstart;
}

// np.interp's result for x in (xj, xj1) (numpy's arr_interp: slope form, then the
// other side when the first gives NaN).
__device__ __forceinline__ double interp_between(double x, double xj, double xj1, double fj,
                                                 double fj1) {
    const double slope = (fj1 - fj) / (xj1 - xj);
    double r = slope * (x - xj) + fj;
    if (r != r) {
        r = slope * (x - xj1) + fj1;
        if (r != r && fj == fj1) r = fj;
    }
    return r;
}

// The same with numpy's precomputed slope of the interval.
__device__ __forceinline__ double interp_slope(double x, double xj, double xj1, double fj,
                                               double fj1, double slope) {
    double r = slope * (x - xj) + fj;
    if (r != r) {
        r = slope * (x - xj1) + fj1;
        if (r != r && fj == fj1) r = fj;
    }
    return r;
}

// Largest i with xp(i) <= x for x in [xp(0), xp(n-1)] (np.interp's search result on a
// non-decreasing grid); -1 below, n above.
template <typename XP>
__device__ __forceinline__ int search(double x, int n, XP xp) {
    if (x > xp(n - 1)) return n;
    if (x < xp(0)) return -1;
    int lo = 0, hi = n;  // invariant: xp(lo) <= x, answer in [lo, hi)
    while (hi - lo > 1) {
        const int mid = lo + ((hi - lo) >> 1);
        if (x >= xp(mid)) lo = mid;
        else hi = mid;
    }
    return lo;
}

// np.interp(x, xp, fp, left, right) for one x.
template <typename XP, typename FP>
__device__ __forceinline__ double interp1(double x, int n, XP xp, FP fp, double left,
                                          double right) {
    if (n == 1) {  // numpy's single-point branch: NaN compares false -> fp[0]
        const double x0 = xp(0);
        return x < x0 ? left : (x > x0 ? right : fp(0));
    }
    if (x != x) return x;
    const int j = search(x, n, xp);
    if (j < 0) return left;
    if (j >= n) return right;
    if (j == n - 1) return fp(j);
    const double xj = xp(j);
    if (xj == x) return fp(j);
    return interp_between(x, xj, xp(j + 1), fp(j), fp(j + 1));
}

// ---- eq_hist bins (np.histogram with uniform bins) --------------------------------------

__device__ __forceinline__ double edge(const Params &P, int k) {
    if (k >= P.nb) return P.last;
    if (P.kind == K_F32) {
        float y = (float)k;
        if (P.step_zero) { y = y / (float)P.nb; y = y * (float)P.delta; }
        else y = y * (float)P.step;
        return (double)(y + (float)P.first);
    }
    double y = (double)k;
    if (P.step_zero) { y = y / (double)P.nb; y = y * P.delta; }
    else y = y * P.step;
    return y + P.first;
}

// Bin of a value inside [first, last] (numpy/lib/_histograms_impl.py histogram, the
// uniform-bins loop: scaled index, last edge into the last bin, +-1 corrections).
__device__ __forceinline__ int bin_of(const Params &P, double x) {
    int64_t k;
    if (P.kind == K_F32) {
        const float a = (float)x;
        const float f = ((a - (float)P.first) / (float)P.denom) * (float)kBins;
        k = (int64_t)f;
    } else {
        const double f = ((x - P.first) / P.denom) * (double)kBins;
        k = (int64_t)f;
    }
    if (k >= P.nb) k = P.nb - 1;
    if (k < 0) k = 0;
    if (x < edge(P, (int)k) && k > 0) k -= 1;
    if (k != P.nb - 1 && x >= edge(P, (int)k + 1)) k += 1;
    return (int)k;
}

// ---- transforms (transfer_functions:225-228) -------------------------------------------

__device__ __forceinline__ double transform(double x, int how, int kind) {
    if (how == DSX_HOW_LOG) {
        if (kind == K_F32) return (double)(float)log1p((double)(float)x);
        return log1p(x);
    }
    if (how == DSX_HOW_CBRT) {
        // float32 ** (1/3.) is a float32 power with exponent float32(1/3)
        if (kind == K_F32) return (double)(float)pow((double)(float)x, (double)(1.0f / 3.0f));
        return pow(x, 1.0 / 3.0);
    }
    return x;
}

// ==== 2D passes =============================================================================

// Pass 1: count / min / max of the unmasked values.
template <typename T, bool BOOLEAN>
__global__ __launch_bounds__(kThreads) void raw_stats_kernel(const T *data, int64_t n, Ws w) {
    Part p;
    part_init(p);
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const T v = data[i];
        if (!masked_raw<T, BOOLEAN>(v)) part_add(p, BOOLEAN ? 1.0 : (double)v);
    }
    block_part(p, w.parts);
}

// Plan A: the offset (nanmin of the unmasked values, or span[0] cast to the dtype) and
// the clip bounds.
__global__ __launch_bounds__(kPlanThreads) void plan_raw_kernel(Ws w, int nparts, int kind,
                                                            dsx_shade_desc d) {
    const Part p = merge_parts(w.parts, nparts);
    if (threadIdx.x != 0) return;
    Params &P = *w.P;
    P.n = p.n; P.mn = p.mn; P.mx = p.mx; P.pmin = p.pmin; P.zero = p.zero;
    P.all_masked = p.n == 0;
    P.kind = kind;
    P.zero_masked = 0;
    P.offset = d.has_span ? cast_to(d.span_lo, kind) : p.mn;
    P.clip_lo = d.span_lo; P.clip_hi = d.span_hi;
    P.clip_lo_set = cast_to(d.span_lo, kind);
    P.clip_hi_set = cast_to(d.span_hi, kind);
}

// Pass 2: the offset values (masked -> NaN): masked_clip_2d (_cpu_utils.py) then
// `data -= offset` in the plane's dtype.
template <typename T, bool BOOLEAN>
__global__ __launch_bounds__(kThreads) void offset_kernel(const T *data, int64_t n, Ws w,
                                                          int has_span) {
    const Params P = *w.P;
    Part p;
    part_init(p);
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const T raw = data[i];
        double out;
        if (masked_raw<T, BOOLEAN>(raw)) {
            out = __builtin_nan("");
        } else {
            double v = BOOLEAN ? 1.0 : (double)raw;
            if (has_span) {
                if (v < P.clip_lo) v = P.clip_lo_set;
                else if (v > P.clip_hi) v = P.clip_hi_set;
            }
            if constexpr (std::is_same<T, float>::value) out = (double)((float)v - (float)P.offset);
            else out = v - P.offset;  // integers: exact below 2^53; float64: one rounding
            part_add(p, out);
        }
        w.val[i] = out;
    }
    block_part(p, w.parts);
}

// Plan B: stats of the offset values and eq_hist's bins (transfer_functions:191-217).
__global__ __launch_bounds__(kPlanThreads) void plan_val_kernel(Ws w, int nparts, int how) {
    const Part p = merge_parts(w.parts, nparts);
    if (threadIdx.x != 0) return;
    Params &P = *w.P;
    P.vn = p.n; P.vmin = p.mn; P.vmax = p.mx;
    P.int_path = 0; P.nb = 0; P.L = 0; P.levels = 0; P.total = 0.0;
    if (how != DSX_HOW_EQ_HIST || p.n == 0) return;
    if (P.kind == K_INT && p.mx - p.mn < (double)kBins) {
        // exact value counts: bins arange(vmin, vmax + 1)
        P.int_path = 1;
        P.nb = (int32_t)(p.mx - p.mn) + 1;
        P.first = p.mn;
        P.last = p.mx;
        return;
    }
    P.nb = kBins;
    double first = p.mn, last = p.mx;
    if (P.kind == K_F32) {
        float f = (float)first, l = (float)last;
        if (f == l) { f = f - 0.5f; l = l + 0.5f; }
        const float delta = l - f;
        const float step = delta / (float)kBins;
        P.first = f; P.last = l; P.delta = delta; P.denom = delta; P.step = step;
        P.step_zero = step == 0.0f;
    } else {
        if (first == last) { first -= 0.5; last += 0.5; }
        const double delta = last - first;
        const double step = delta / (double)kBins;
        P.first = first; P.last = last; P.delta = delta; P.denom = delta; P.step = step;
        P.step_zero = step == 0.0;
    }
}

__global__ __launch_bounds__(kThreads) void zero_hist_kernel(Ws w) {
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < kBins; i += gridDim.x * kThreads)
        w.hist[i] = 0u;
}

// Histogram of the valid offset values: each workgroup counts its cells into one half
// of the bins in LDS, then adds its non-empty bins to the global histogram.
__global__ __launch_bounds__(1024) void hist_kernel(int64_t n, Ws w) {
    __shared__ uint32_t h[kHalfBins];
    const Params &P = *w.P;
    if (P.vn == 0 || P.nb == 0) return;
    const int b0 = blockIdx.y * kHalfBins;
    if (b0 >= P.nb) return;
    for (int k = threadIdx.x; k < kHalfBins; k += 1024) h[k] = 0u;
    __syncthreads();
    const int64_t c0 = (int64_t)blockIdx.x * kHistCells;
    const int64_t c1 = c0 + kHistCells < n ? c0 + kHistCells : n;
    for (int64_t i = c0 + threadIdx.x; i < c1; i += 1024) {
        const double x = w.val[i];
        if (x != x) continue;
        const int k = P.int_path ? (int)(x - P.first) : bin_of(P, x);
        const int r = k - b0;
        if (r >= 0 && r < kHalfBins) atomicAdd(&h[r], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kHalfBins; k += 1024) {
        const uint32_t c = h[k];
        if (c) atomicAdd(&w.hist[b0 + k], c);
    }
}

// Inclusive scan over a 1024-thread workgroup (wave scans, then the 16 wave totals).
template <typename T>
__device__ __forceinline__ T block_scan(T v, T *wave_tot, T &total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const T y = __shfl_up(v, off, 64);
        if (lane >= off) v += y;
    }
    if (lane == 63) wave_tot[wid] = v;
    __syncthreads();
    if (wid == 0) {
        T t = lane < 16 ? wave_tot[lane] : (T)0;
        for (int off = 1; off < 16; off <<= 1) {
            const T y = __shfl_up(t, off, 64);
            if (lane >= off) t += y;
        }
        if (lane < 16) wave_tot[lane] = t;
    }
    __syncthreads();
    const T pre = wid ? wave_tot[wid - 1] : (T)0;
    total = wave_tot[15];
    __syncthreads();
    return v + pre;
}

// Per 1024-bin chunk of the histogram: number of values and of non-empty bins.
__global__ __launch_bounds__(kCdfThreads) void chunk_kernel(Ws w) {
    __shared__ unsigned long long ws[16];
    __shared__ uint32_t wn[16];
    const Params &P = *w.P;
    if (P.vn == 0 || P.nb == 0) return;
    const int k = blockIdx.x * kCdfThreads + threadIdx.x;
    const uint32_t c = k < P.nb ? w.hist[k] : 0u;
    unsigned long long st;
    uint32_t nt;
    block_scan<unsigned long long>(c, ws, st);
    block_scan<uint32_t>(c != 0u, wn, nt);
    if (threadIdx.x == 0) { w.chunk_sum[blockIdx.x] = st; w.chunk_nz[blockIdx.x] = nt; }
}

// cumsum / total of the histogram (eq_hist: `cdf = hist.cumsum(); cdf /= cdf[-1]`); the
// float path keeps only the non-empty bins, their centres and each one's rank.
__global__ __launch_bounds__(kCdfThreads) void cdf_kernel(Ws w) {
    __shared__ unsigned long long ws[16];
    __shared__ uint32_t wn[16];
    __shared__ unsigned long long pre_s, tot_s;
    __shared__ uint32_t pre_n, tot_n;
    Params &P = *w.P;
    const int nb = P.nb;
    if (P.vn == 0 || nb == 0) return;
    const int b = blockIdx.x;
    if (threadIdx.x == 0) {
        unsigned long long ps = 0, ts = 0;
        uint32_t pn = 0, tn = 0;
        for (int i = 0; i < kCdfBlocks; ++i) {
            if (i < b) { ps += w.chunk_sum[i]; pn += w.chunk_nz[i]; }
            ts += w.chunk_sum[i];
            tn += w.chunk_nz[i];
        }
        pre_s = ps; tot_s = ts; pre_n = pn; tot_n = tn;
    }
    __syncthreads();
    const int k = b * kCdfThreads + threadIdx.x;
    const uint32_t c = k < nb ? w.hist[k] : 0u;
    unsigned long long st;
    uint32_t nt;
    const unsigned long long cum = pre_s + block_scan<unsigned long long>(c, ws, st);
    const uint32_t pos = pre_n + block_scan<uint32_t>(c != 0u, wn, nt);  // inclusive
    const double dt = (double)tot_s;
    if (k < nb) {
        if (P.int_path) {
            w.cdf[k] = (double)cum / dt;
        } else if (c) {
            double centre;
            if (P.kind == K_F32) centre = (double)(((float)edge(P, k) + (float)edge(P, k + 1)) / 2.0f);
            else centre = (edge(P, k) + edge(P, k + 1)) / 2.0;
            w.xp[pos - 1] = centre;
            w.cdf[pos - 1] = (double)cum / dt;
            w.rank[k] = (int32_t)pos - 1;
        }
    }
    if (b == 0 && threadIdx.x == 0) {
        P.levels = (int32_t)tot_n;
        P.L = P.int_path ? nb : (int32_t)tot_n;
        P.total = dt;
    }
}

// Slopes between consecutive (centre, cdf) points, as np.interp precomputes them.
__global__ __launch_bounds__(kThreads) void eslope_kernel(Ws w) {
    const Params &P = *w.P;
    if (P.vn == 0 || P.int_path || P.nb == 0) return;
    const int j = blockIdx.x * kThreads + threadIdx.x;
    if (j + 1 < P.L) w.eslope[j] = (w.cdf[j + 1] - w.cdf[j]) / (w.xp[j + 1] - w.xp[j]);
}

// Pass 3: the transformed values (in place) and their min / max.
__global__ __launch_bounds__(kThreads) void transform_kernel(int64_t n, Ws w, int how) {
    const Params P = *w.P;
    const double *xp = w.xp, *cdf = w.cdf;
    Part p;
    part_init(p);
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const double x = w.val[i];
        if (x != x) continue;
        double t;
        if (how == DSX_HOW_EQ_HIST) {
            if (P.int_path) {
                t = cdf[(int64_t)(x - P.first)];
            } else {
                // x lies in its own (non-empty) bin k: the largest centre <= x is bin k's
                // or the previous non-empty bin's -- np.interp's search result in O(1).
                const int L = P.L;
                const int r = w.rank[bin_of(P, x)];
                int j = x >= xp[r] ? r : r - 1;
                if (!((j < 0 || xp[j] <= x) && (j + 1 >= L || xp[j + 1] > x)))
                    j = search(x, L, [&](int q) { return xp[q]; });
                if (L == 1 || j < 0) t = cdf[0];
                else if (j >= L - 1) t = cdf[L - 1];
                else if (xp[j] == x) t = cdf[j];
                else t = interp_slope(x, xp[j], xp[j + 1], cdf[j], cdf[j + 1], w.eslope[j]);
            }
        } else {
            t = transform(x, how, P.kind);
        }
        w.val[i] = t;
        if (t == t) part_add(p, t);
    }
    block_part(p, w.parts);
}

// Plan C: the span of the transformed values (or of the transformed given span) and
// the knots np.linspace(span[0], span[1], nk) (transfer_functions:293-311, 318-332).
__global__ __launch_bounds__(kPlanThreads) void plan_span_kernel(Ws w, int nparts, dsx_shade_desc d,
                                                             int nk, int knots_f64) {
    const Part p = merge_parts(w.parts, nparts);
    __shared__ Params sP;
    if (threadIdx.x == 0) {
        Params &P = sP;
        P = *w.P;
        P.tmin = p.mn; P.tmax = p.mx;
        double s0, s1;
        int f32;
        if (d.has_span) {
            s0 = transform(0.0, d.how, K_F64);
            s1 = transform(d.span_delta, d.how, K_F64);
            f32 = 0;
        } else {
            s0 = p.mn;
            s1 = p.mx;
            // nanmin / nanmax of a float32 plane are float32 scalars: linspace stays float32
            f32 = !knots_f64 && P.kind == K_F32 && d.how != DSX_HOW_EQ_HIST;
            if (d.rescale_discrete_levels && d.how == DSX_HOW_EQ_HIST) {
                // _rescale_discrete_levels (transfer_functions:236-248)
                const double m = -0.5 / 98.0;
                const double c = 1.5 - 2.0 * m;
                const double multiple = m * (double)P.levels + c;
                if (multiple > 1.0) {
                    const double lower = s1 - multiple * (s1 - s0);
                    s0 = lower > 0.0 ? lower : 0.0;
                    s1 = 1.0;
                }
            }
        }
        P.kf32 = f32;
        P.nk = nk;
        if (f32) {
            const float a = (float)s0, b = (float)s1;
            const float delta = b - a;
            P.kstart = a; P.kstop = b; P.kdelta = delta;
            if (nk > 1) {
                const float step = delta / (float)(nk - 1);
                P.kstep = step;
                P.kmode = step == 0.0f ? 1 : 0;
            } else {
                P.kmode = 2;
            }
        } else {
            const double delta = s1 - s0;
            P.kstart = s0; P.kstop = s1; P.kdelta = delta;
            if (nk > 1) {
                const double step = delta / (double)(nk - 1);
                P.kstep = step;
                P.kmode = step == 0.0 ? 1 : 0;
            } else {
                P.kmode = 2;
            }
        }
        *w.P = P;
    }
    __syncthreads();
    const Params &Q = sP;
    // knots and the slopes np.interp precomputes: colour channels (list colormap), the
    // alpha steps (one colour) or the two-point alpha ramp (categorical)
    if (d.kind == DSX_SHADE_CATEGORICAL) {
        if (threadIdx.x == 0) {
            w.kx[0] = Q.kstart;
            w.kx[1] = Q.kstop;
            w.ks[0] = (d.alpha - d.min_alpha) / (Q.kstop - Q.kstart);
        }
        return;
    }
    const double d1 = (d.min_alpha + 1.0) - d.min_alpha;
    auto fa = [&](int k) {
        return k == 0 ? d.min_alpha : (k == 1 ? d.min_alpha + 1.0 : d.min_alpha + (double)k * d1);
    };
    for (int k = threadIdx.x; k < Q.nk; k += kPlanThreads) {
        const double xk = knot(Q, k);
        w.kx[k] = xk;
        if (k + 1 < Q.nk) {
            const double dx = knot(Q, k + 1) - xk;
            if (d.kind == DSX_SHADE_CMAP) {
                for (int c = 0; c < 3; ++c)
                    w.ks[3 * k + c] = ((double)d.colors[3 * (k + 1) + c] - (double)d.colors[3 * k + c]) / dx;
            } else {
                w.ks[k] = (fa(k + 1) - fa(k)) / dx;
            }
        }
    }
}

__device__ __forceinline__ uint32_t pack(uint32_t r, uint32_t g, uint32_t b, uint32_t a) {
    return r | (g << 8) | (b << 16) | (a << 24);
}

// Pass 4 (list colormap): r, g, b = interp(t, knots, channel, left=255), alpha constant.
__global__ __launch_bounds__(kThreads) void cmap_kernel(int64_t n, Ws w, const uint8_t *colors,
                                                        double alpha, uint32_t *out) {
    const int nk = w.P->nk;
    const int all_masked = w.P->all_masked;
    const double *kx = w.kx, *ks = w.ks;
    const uint32_t a_const = u8_of(alpha);
    auto xp = [&](int k) { return kx[k]; };
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        if (all_masked) { out[i] = 0u; continue; }
        const double t = w.val[i];
        uint32_t ch[3];
        if (nk == 1) {  // numpy's single-point branch: NaN -> the colour
            const double x0 = kx[0];
            for (int c = 0; c < 3; ++c) ch[c] = t < x0 ? 255u : colors[c];
        } else if (t != t) {
            ch[0] = ch[1] = ch[2] = 0u;  // nan_to_num(NaN) = 0
        } else {
            const int j = search(t, nk, xp);
            if (j < 0) {
                ch[0] = ch[1] = ch[2] = 255u;
            } else if (j >= nk - 1) {
                for (int c = 0; c < 3; ++c) ch[c] = colors[3 * (nk - 1) + c];
            } else {
                const double xj = kx[j];
                if (xj == t) {
                    for (int c = 0; c < 3; ++c) ch[c] = colors[3 * j + c];
                } else {
                    const double xj1 = kx[j + 1];
                    for (int c = 0; c < 3; ++c)
                        ch[c] = u8_of(interp_slope(t, xj, xj1, (double)colors[3 * j + c],
                                                   (double)colors[3 * (j + 1) + c],
                                                   ks[3 * j + c]));
                }
            }
        }
        out[i] = pack(ch[0], ch[1], ch[2], t != t ? 0u : a_const);
    }
}

// Pass 4 (one colour): rgb constant, alpha = interp(t, knots, arange(min_alpha, alpha+1),
// left=0, right=255).
__global__ __launch_bounds__(kThreads) void color_kernel(int64_t n, Ws w, const uint8_t *colors,
                                                         double min_alpha, uint32_t *out) {
    const int nk = w.P->nk;
    const int all_masked = w.P->all_masked;
    const double *kx = w.kx, *ks = w.ks;
    const uint32_t rgb = pack(colors[0], colors[1], colors[2], 0u);
    // np.arange(min_alpha, alpha + 1): start, start + 1, then start + k * (second - start)
    const double d1 = (min_alpha + 1.0) - min_alpha;
    auto fp = [&](int k) {
        return k == 0 ? min_alpha : (k == 1 ? min_alpha + 1.0 : min_alpha + (double)k * d1);
    };
    auto xp = [&](int k) { return kx[k]; };
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        if (all_masked) { out[i] = 0u; continue; }
        const double t = w.val[i];
        double a;
        if (nk == 1) {
            a = t < kx[0] ? 0.0 : (t > kx[0] ? 255.0 : fp(0));
        } else if (t != t) {
            a = 0.0;
        } else {
            const int j = search(t, nk, xp);
            if (j < 0) a = 0.0;
            else if (j >= nk) a = 255.0;
            else if (j == nk - 1 || kx[j] == t) a = fp(j);
            else a = interp_slope(t, kx[j], kx[j + 1], fp(j), fp(j + 1), ks[j]);
        }
        if (a != a) a = 0.0;
        out[i] = rgb | (u8_of(a) << 24);
    }
}

// ==== 3D (categorical) passes ==================================================================

template <typename T>
__device__ __forceinline__ bool is_nan_v(T v) {
    if constexpr (std::is_floating_point<T>::value) return v != v;
    else return false;
}

// nanmin over every value (the default color_baseline, transfer_functions:393-396).
template <typename T>
__global__ __launch_bounds__(kThreads) void cat_min_kernel(const T *data, int64_t n, Ws w) {
    Part p;
    part_init(p);
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const T v = data[i];
        if (!is_nan_v(v)) part_add(p, (double)v);
    }
    block_part(p, w.parts);
}

__global__ __launch_bounds__(kPlanThreads) void plan_baseline_kernel(Ws w, int nparts,
                                                                 dsx_shade_desc d, int kind) {
    const Part p = merge_parts(w.parts, nparts);
    if (threadIdx.x != 0) return;
    Params &P = *w.P;
    // nanmin of an all-NaN plane is NaN: neither branch subtracts anything
    P.offset = d.has_baseline ? d.baseline : (p.n ? p.mn : __builtin_nan(""));
    P.kind = kind;
}

// Per pixel: the colour mix (transfer_functions:398-441, in float32 as numpy computes it)
// and the total over categories (nansum_missing, utils.py:161-181) for the alpha.
template <typename T>
__global__ __launch_bounds__(kThreads) void cat_pixel_kernel(const T *data, int64_t npix,
                                                             int ncat, const uint8_t *colors,
                                                             int has_baseline, Ws w,
                                                             uint32_t *out) {
    const double base = w.P->offset;
    Part p;
    part_init(p);
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < npix;
         i += (int64_t)gridDim.x * kThreads) {
        const T *row = data + i * (int64_t)ncat;
        float ctot = 0.0f, sr = 0.0f, sg = 0.0f, sb = 0.0f;
        float mcnt = 0.0f, mr = 0.0f, mg = 0.0f, mb = 0.0f;
        // the total keeps the sum dtype: uint64 / int64 for integers, the float type else
        unsigned long long utot = 0;
        long long itot = 0;
        float ftot = 0.0f;
        double dtot = 0.0;
        int valid = 0;
        for (int c = 0; c < ncat; ++c) {
            const T v = row[c];
            const float cr = colors[3 * c], cg = colors[3 * c + 1], cb = colors[3 * c + 2];
            if (is_nan_v(v)) continue;  // NaN -> colour weight 0, not present
            ++valid;
            mcnt += 1.0f; mr += cr; mg += cg; mb += cb;
            float cd;
            if constexpr (std::is_same<T, float>::value) {
                float x = v;
                if (base > 0.0) x = x - (float)base;
                else if (base == base) x = x + (float)(-base);
                if (has_baseline && x < 0.0f) x = 0.0f;
                cd = x;
                ftot += v;
            } else if constexpr (std::is_same<T, double>::value) {
                double x = v;
                if (base > 0.0) x = x - base;
                else if (base == base) x = x + (-base);
                if (has_baseline && x < 0.0) x = 0.0;
                cd = (float)x;
                dtot += v;
            } else if constexpr (std::is_same<T, uint32_t>::value) {
                uint32_t x = v;
                if (base > 0.0) x = x - (uint32_t)(long long)base;       // wraps like numpy
                else if (base == base) x = x + (uint32_t)(long long)(-base);
                cd = (float)x;
                utot += v;
            } else {
                long long x = (long long)v;
                if (base > 0.0) x = x - (long long)base;
                else if (base == base) x = x + (long long)(-base);
                x = (long long)(T)x;  // the array keeps its own width
                if (has_baseline && x < 0) x = 0;
                cd = (float)x;
                itot += (long long)v;
            }
            ctot += cd;
            sr += cd * cr; sg += cd * cg; sb += cd * cb;
        }
        uint32_t r, g, b;
        if (ctot == 0.0f) {
            r = u8_of((double)(mr / mcnt)); g = u8_of((double)(mg / mcnt));
            b = u8_of((double)(mb / mcnt));
        } else {
            r = u8_of((double)(sr / ctot)); g = u8_of((double)(sg / ctot));
            b = u8_of((double)(sb / ctot));
        }
        out[i] = pack(r, g, b, 0u);
        double tot;
        if (valid == 0) tot = __builtin_nan("");
        else if constexpr (std::is_same<T, float>::value) tot = (double)ftot;
        else if constexpr (std::is_same<T, double>::value) tot = dtot;
        else if constexpr (std::is_same<T, uint32_t>::value) tot = (double)utot;
        else tot = (double)itot;
        w.val[i] = tot;
        if (tot == tot) part_add(p, tot);
    }
    block_part(p, w.parts);
}

// Plan for the alpha of a categorical plane (_interpolate_alpha, transfer_functions:
// 474-518): offset, the masking of zero totals of unsigned planes, clip bounds.
__global__ __launch_bounds__(kPlanThreads) void plan_total_kernel(Ws w, int nparts, dsx_shade_desc d,
                                                              int data_kind, int unsigned_data) {
    const Part p = merge_parts(w.parts, nparts);
    if (threadIdx.x != 0) return;
    Params &P = *w.P;
    P.n = p.n; P.mn = p.mn; P.mx = p.mx; P.pmin = p.pmin; P.zero = p.zero;
    P.all_masked = 0;
    P.zero_masked = 0;
    int kind = data_kind;
    if (!d.has_span) {
        P.offset = p.mn;
        if (unsigned_data && p.n && p.mn == 0.0) {
            P.zero_masked = 1;
            kind = K_F64;  // np.where(~mask, total, np.nan) is float64
            if (p.pmin < __builtin_inf()) P.offset = p.pmin;
        }
    } else {
        P.offset = cast_to(d.span_lo, data_kind);
        if (unsigned_data && p.n && p.mn == 0.0) {
            P.zero_masked = 1;
            kind = K_F64;
        }
    }
    P.kind = kind;
    P.clip_lo = d.span_lo; P.clip_hi = d.span_hi;
    P.clip_lo_set = cast_to(d.span_lo, kind);
    P.clip_hi_set = cast_to(d.span_hi, kind);
}

// Offset values of the totals (in place), masked -> NaN.
__global__ __launch_bounds__(kThreads) void total_offset_kernel(int64_t npix, Ws w, int has_span) {
    const Params P = *w.P;
    Part p;
    part_init(p);
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < npix;
         i += (int64_t)gridDim.x * kThreads) {
        double v = w.val[i];
        if (v != v || (P.zero_masked && v == 0.0)) {
            w.val[i] = __builtin_nan("");
            continue;
        }
        if (has_span) {
            if (v < P.clip_lo) v = P.clip_lo_set;
            else if (v > P.clip_hi) v = P.clip_hi_set;
        }
        double out;
        if (P.kind == K_F32) out = (double)((float)v - (float)P.offset);
        else out = v - P.offset;
        w.val[i] = out;
        part_add(p, out);
    }
    block_part(p, w.parts);
}

// Alpha of a categorical pixel: interp(a_scaled, norm_span, [min_alpha, alpha], 0, 255).
__global__ __launch_bounds__(kThreads) void cat_alpha_kernel(int64_t npix, Ws w, double min_alpha,
                                                             double alpha, uint32_t *out) {
    const double x0 = w.kx[0], x1 = w.kx[1], slope = w.ks[0];
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < npix;
         i += (int64_t)gridDim.x * kThreads) {
        const double t = w.val[i];
        double a;
        if (t != t) a = 0.0;
        else if (t > x1) a = 255.0;
        else if (t < x0) a = 0.0;
        else if (t >= x1) a = alpha;        // search result j == 1 (the last point)
        else if (t == x0) a = min_alpha;
        else a = interp_slope(t, x0, x1, min_alpha, alpha, slope);
        if (a != a) a = 0.0;
        out[i] = (out[i] & 0x00FFFFFFu) | (u8_of(a) << 24);
    }
}

// ==== 2D colour key (_apply_discrete_colorkey, transfer_functions:541-613) =======================

template <typename T>
__global__ __launch_bounds__(kThreads) void color_key_kernel(const T *data, int64_t n,
                                                             const double *keys,
                                                             const uint8_t *colors, int nkeys,
                                                             int black, double alpha,
                                                             uint32_t *out) {
    const uint32_t a = u8_of(alpha) << 24;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const double v = (double)data[i];
        int hit = -1;
        for (int k = 0; k < nkeys; ++k)
            if (keys[k] == v) hit = k;  // later keys overwrite earlier ones
        if (hit < 0) { out[i] = 0u; continue; }
        out[i] = (black ? 0u : pack(colors[3 * hit], colors[3 * hit + 1], colors[3 * hit + 2], 0u)) | a;
    }
}

// ==== host side ===================================================================================

int grid_parts(int64_t n) {
    int64_t g = (n + kThreads * 4 - 1) / (kThreads * 4);
    if (g < 1) g = 1;
    if (g > kMaxParts) g = kMaxParts;
    return (int)g;
}

int check(const dsx_shade_desc *d) {
    if (!d) return DSX_ERR_NULL;
    if (d->H < 1 || d->W < 1 || d->ncat < 1) return DSX_ERR_SHAPE;
    if (d->H * d->W > ((int64_t)1 << 40)) return DSX_ERR_SHAPE;
    if (!d->data || !d->out) return DSX_ERR_NULL;
    switch (d->dtype) {
    case DSX_U32: case DSX_BOOL: case DSX_I32: case DSX_I64: case DSX_F32: case DSX_F64: break;
    default: return DSX_ERR_DTYPE;
    }
    if (d->how < DSX_HOW_LINEAR || d->how > DSX_HOW_EQ_HIST) return DSX_ERR_MODE;
    switch (d->kind) {
    case DSX_SHADE_CMAP:
        if (!d->colors) return DSX_ERR_NULL;
        if (d->ncat != 1 || d->ncolors < 1 || d->ncolors > kMaxKnots) return DSX_ERR_SHAPE;
        break;
    case DSX_SHADE_COLOR:
        if (d->ncat != 1 || d->ncolors != 1 || d->nalpha < 1 || d->nalpha > kMaxKnots)
            return DSX_ERR_SHAPE;
        if (!d->colors) return DSX_ERR_NULL;
        break;
    case DSX_SHADE_CATEGORICAL:
        if (d->ncolors != d->ncat || d->dtype == DSX_BOOL) return DSX_ERR_SHAPE;
        if (!d->colors) return DSX_ERR_NULL;
        break;
    case DSX_SHADE_COLOR_KEY:
        if (d->ncat != 1 || d->ncolors < 0) return DSX_ERR_SHAPE;
        if (d->ncolors > 0 && (!d->colors || !d->keys)) return DSX_ERR_NULL;
        break;
    default: return DSX_ERR_MODE;
    }
    if (d->has_span && d->how == DSX_HOW_EQ_HIST) return DSX_ERR_MODE;
    return DSX_OK;
}

int data_kind(int dtype) {
    return dtype == DSX_F32 ? K_F32 : (dtype == DSX_F64 ? K_F64 : K_INT);
}

// The shared tail: eq_hist (when asked), transform, span + knots.
int tail(const dsx_shade_desc *d, Ws w, int64_t npix, int nk, int knots_f64, hipStream_t st) {
    const int g = grid_parts(npix);
    int s;
    hipLaunchKernelGGL(plan_val_kernel, dim3(1), dim3(kPlanThreads), 0, st, w, g, d->how);
    if ((s = launch_status())) return s;
    if (d->how == DSX_HOW_EQ_HIST) {
        hipLaunchKernelGGL(zero_hist_kernel, dim3(kBins / kThreads), dim3(kThreads), 0, st, w);
        if ((s = launch_status())) return s;
        const int64_t hb = (npix + kHistCells - 1) / kHistCells;
        hipLaunchKernelGGL(hist_kernel, dim3((unsigned)hb, 2), dim3(1024), 0, st, npix, w);
        if ((s = launch_status())) return s;
        hipLaunchKernelGGL(chunk_kernel, dim3(kCdfBlocks), dim3(kCdfThreads), 0, st, w);
        if ((s = launch_status())) return s;
        hipLaunchKernelGGL(cdf_kernel, dim3(kCdfBlocks), dim3(kCdfThreads), 0, st, w);
        if ((s = launch_status())) return s;
        hipLaunchKernelGGL(eslope_kernel, dim3(kBins / kThreads), dim3(kThreads), 0, st, w);
        if ((s = launch_status())) return s;
    }
    hipLaunchKernelGGL(transform_kernel, dim3(g), dim3(kThreads), 0, st, npix, w, d->how);
    if ((s = launch_status())) return s;
    hipLaunchKernelGGL(plan_span_kernel, dim3(1), dim3(kPlanThreads), 0, st, w, g, *d, nk, knots_f64);
    return launch_status();
}

template <typename T, bool BOOLEAN>
int run2d(const dsx_shade_desc *d, Ws w, hipStream_t st) {
    const int64_t n = d->H * d->W;
    const T *data = reinterpret_cast<const T *>(d->data);
    const int g = grid_parts(n);
    int s;
    hipLaunchKernelGGL((raw_stats_kernel<T, BOOLEAN>), dim3(g), dim3(kThreads), 0, st, data, n, w);
    if ((s = launch_status())) return s;
    hipLaunchKernelGGL(plan_raw_kernel, dim3(1), dim3(kPlanThreads), 0, st, w, g,
                       data_kind(d->dtype), *d);
    if ((s = launch_status())) return s;
    hipLaunchKernelGGL((offset_kernel<T, BOOLEAN>), dim3(g), dim3(kThreads), 0, st, data, n, w,
                       d->has_span);
    if ((s = launch_status())) return s;
    const int nk = d->kind == DSX_SHADE_CMAP ? d->ncolors : d->nalpha;
    if ((s = tail(d, w, n, nk, 0, st))) return s;
    if (d->kind == DSX_SHADE_CMAP)
        hipLaunchKernelGGL(cmap_kernel, dim3(g), dim3(kThreads), 0, st, n, w, d->colors, d->alpha,
                           d->out);
    else
        hipLaunchKernelGGL(color_kernel, dim3(g), dim3(kThreads), 0, st, n, w, d->colors,
                           d->min_alpha, d->out);
    return launch_status();
}

template <typename T>
int run3d(const dsx_shade_desc *d, Ws w, hipStream_t st) {
    const int64_t npix = d->H * d->W;
    const int64_t n = npix * d->ncat;
    const T *data = reinterpret_cast<const T *>(d->data);
    int s;
    const int gn = grid_parts(n), gp = grid_parts(npix);
    hipLaunchKernelGGL((cat_min_kernel<T>), dim3(gn), dim3(kThreads), 0, st, data, n, w);
    if ((s = launch_status())) return s;
    hipLaunchKernelGGL(plan_baseline_kernel, dim3(1), dim3(kPlanThreads), 0, st, w, gn, *d,
                       data_kind(d->dtype));
    if ((s = launch_status())) return s;
    hipLaunchKernelGGL((cat_pixel_kernel<T>), dim3(gp), dim3(kThreads), 0, st, data, npix,
                       (int)d->ncat, d->colors, d->has_baseline, w, d->out);
    if ((s = launch_status())) return s;
    hipLaunchKernelGGL(plan_total_kernel, dim3(1), dim3(kPlanThreads), 0, st, w, gp, *d,
                       data_kind(d->dtype), (int)(d->dtype == DSX_U32));
    if ((s = launch_status())) return s;
    hipLaunchKernelGGL(total_offset_kernel, dim3(gp), dim3(kThreads), 0, st, npix, w, d->has_span);
    if ((s = launch_status())) return s;
    if ((s = tail(d, w, npix, 2, 1, st))) return s;
    hipLaunchKernelGGL(cat_alpha_kernel, dim3(gp), dim3(kThreads), 0, st, npix, w, d->min_alpha,
                       d->alpha, d->out);
    return launch_status();
}

template <typename T>
int run_key(const dsx_shade_desc *d, hipStream_t st) {
    const int64_t n = d->H * d->W;
    // color_data after the baseline step (transfer_functions:580-592): 1 - b, 1 + |b|
    // or 1, clipped at 0 when a baseline is given; a non-zero value leaves the pixels black.
    double v = 1.0;
    if (d->has_baseline) {
        if (d->baseline > 0) v = 1.0 - d->baseline;
        else if (d->baseline < 0) v = 1.0 + (-d->baseline);
        if (v < 0) v = 0;
    }
    const int black = d->has_baseline && v != 0.0;
    hipLaunchKernelGGL((color_key_kernel<T>), dim3(grid_parts(n)), dim3(kThreads), 0, st,
                       reinterpret_cast<const T *>(d->data), n, d->keys, d->colors, d->ncolors,
                       black, d->alpha, d->out);
    return launch_status();
}

}  // namespace
}  // namespace dsx

using namespace dsx;

extern "C" size_t dsx_shade_workspace_size(const dsx_shade_desc *d) {
    if (check(d)) return 0;
    return layout_of(d->H * d->W).total;
}

extern "C" int dsx_shade(const dsx_shade_desc *d, void *workspace, size_t workspace_bytes,
                         void *stream) {
    int s = check(d);
    if (s) return s;
    hipStream_t st = as_stream(stream);
    if (d->kind == DSX_SHADE_COLOR_KEY) {
        switch (d->dtype) {
        case DSX_U32: return run_key<uint32_t>(d, st);
        case DSX_BOOL: return run_key<uint8_t>(d, st);
        case DSX_I32: return run_key<int32_t>(d, st);
        case DSX_I64: return run_key<int64_t>(d, st);
        case DSX_F32: return run_key<float>(d, st);
        default: return run_key<double>(d, st);
        }
    }
    const Layout L = layout_of(d->H * d->W);
    if (!workspace || workspace_bytes < L.total) return DSX_ERR_WORKSPACE;
    uint8_t *ws = reinterpret_cast<uint8_t *>(workspace);
    Ws w;
    w.P = reinterpret_cast<Params *>(ws + L.o_params);
    w.parts = reinterpret_cast<Part *>(ws + L.o_parts);
    w.hist = reinterpret_cast<uint32_t *>(ws + L.o_hist);
    w.xp = reinterpret_cast<double *>(ws + L.o_xp);
    w.cdf = reinterpret_cast<double *>(ws + L.o_cdf);
    w.rank = reinterpret_cast<int32_t *>(ws + L.o_rank);
    w.chunk_sum = reinterpret_cast<unsigned long long *>(ws + L.o_chunk);
    w.chunk_nz = reinterpret_cast<uint32_t *>(ws + L.o_chunk + 8 * kCdfBlocks);
    w.eslope = reinterpret_cast<double *>(ws + L.o_eslope);
    w.kx = reinterpret_cast<double *>(ws + L.o_kx);
    w.ks = reinterpret_cast<double *>(ws + L.o_ks);
    w.val = reinterpret_cast<double *>(ws + L.o_val);
    if (d->kind == DSX_SHADE_CATEGORICAL) {
        switch (d->dtype) {
        case DSX_U32: return run3d<uint32_t>(d, w, st);
        case DSX_I32: return run3d<int32_t>(d, w, st);
        case DSX_I64: return run3d<int64_t>(d, w, st);
        case DSX_F32: return run3d<float>(d, w, st);
        default: return run3d<double>(d, w, st);
        }
    }
    switch (d->dtype) {
    case DSX_U32: return run2d<uint32_t, false>(d, w, st);
    case DSX_BOOL: return run2d<uint8_t, true>(d, w, st);
    case DSX_I32: return run2d<int32_t, false>(d, w, st);
    case DSX_I64: return run2d<int64_t, false>(d, w, st);
    case DSX_F32: return run2d<float, false>(d, w, st);
    default: return run2d<double, false>(d, w, st);
    }
}
