// dsx_common.h -- shared device helpers of libdsx (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dsx.h"

namespace dsx {

// ---- status plumbing ---------------------------------------------------------------
void record_hip_error(hipError_t e);
int launch_status();  // DSX_OK or DSX_ERR_HIP after a launch

static inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// ---- dtype loads -------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ double load_f(const void *p, int64_t i) {
    return (double)(reinterpret_cast<const T *>(p)[i]);
}

// Category code of row i for a runtime dtype (the branch is wave-uniform).
__device__ __forceinline__ int64_t load_code(const void *p, int32_t dt, int64_t i) {
    switch (dt) {
    case DSX_I8: return reinterpret_cast<const int8_t *>(p)[i];
    case DSX_U8: return reinterpret_cast<const uint8_t *>(p)[i];
    case DSX_I16: return reinterpret_cast<const int16_t *>(p)[i];
    case DSX_I32: return reinterpret_cast<const int32_t *>(p)[i];
    default: return reinterpret_cast<const int64_t *>(p)[i];
    }
}

// by()'s categorizer (reductions.py:143-259): codes, modulo or binning of one column.
struct Cat {
    int32_t kind, dtype, ncat, nbins, bin_under, bin_over;
    const void *col;
    int64_t offset, modulo;
    double bin0, binsize;
};

inline Cat cat_of(const dsx_categorizer &c) {
    Cat k;
    k.kind = c.kind; k.dtype = c.dtype; k.ncat = c.kind == DSX_CAT_NONE ? 1 : c.ncat;
    k.nbins = c.nbins; k.bin_under = c.bin_under; k.bin_over = c.bin_over;
    k.col = c.col; k.offset = c.offset; k.modulo = c.modulo;
    k.bin0 = c.bin0; k.binsize = c.binsize;
    return k;
}

// A column value read by a pass that reads each row exactly once.  The load is marked
// non-temporal, so the column streams past the memory-side cache instead of evicting the
// workspace and planes the call writes (spill lists, partials), which then stay cached
// from the kernel that writes them to the kernel that reads them: on the 100M-row
// count+mean workload this took the scan from 323 to 262 us.
template <typename T>
__device__ __forceinline__ T stream_load(const T *p) {
    return __builtin_nontemporal_load(p);
}

// The categorizer's input for row i as 64 bits: the integer code, or the bits of the
// value as a double for a float column (exact for float32).
__device__ __forceinline__ int64_t load_cat_raw(const Cat &c, int64_t i) {
    switch (c.dtype) {
    case DSX_F32: return __double_as_longlong((double)reinterpret_cast<const float *>(c.col)[i]);
    case DSX_F64: return __double_as_longlong(reinterpret_cast<const double *>(c.col)[i]);
    default: return load_code(c.col, c.dtype, i);
    }
}

// Category of a row from its raw input (load_cat_raw), in [0, ncat), or -1 when the
// row has no category inside the aggregate.  Registers only: no memory access.
__device__ __forceinline__ int64_t category_raw(const Cat &c, int64_t raw) {
    if (c.kind == DSX_CAT_NONE) return 0;
    int64_t k;
    if (c.kind == DSX_CAT_BINNING) {
        // category_binning.apply (reductions.py:242-258): numpy evaluates
        // (values - bin0) / binsize in the column's float type (float32 stays float32).
        bool nan_v;
        int64_t idx = 0;
        if (c.dtype == DSX_F32) {
            const float v = (float)__longlong_as_double(raw);
            nan_v = v != v;
            if (!nan_v) idx = (int64_t)((v - (float)c.bin0) / (float)c.binsize);
        } else {
            const double v = c.dtype == DSX_F64 ? __longlong_as_double(raw) : (double)raw;
            nan_v = v != v;
            if (!nan_v) idx = (int64_t)((v - c.bin0) / c.binsize);
        }
        if (nan_v) k = c.nbins;
        else if (idx < 0) k = c.bin_under;
        else if (idx >= c.nbins) k = c.bin_over;
        else k = idx;
    } else if (c.kind == DSX_CAT_MODULO) {
        int64_t r = (raw - c.offset) % c.modulo;  // Python's floor modulo
        if (r != 0 && ((r < 0) != (c.modulo < 0))) r += c.modulo;
        k = r;
    } else {
        k = raw;
    }
    if (k < 0) k += c.ncat;  // numba's negative-index wraparound (code -1 = NaN category)
    if (k < 0 || k >= c.ncat) return -1;
    return k;
}

// by()'s categorizer (reductions.py:143-259) applied to row i.
__device__ __forceinline__ int64_t category(const Cat &c, int64_t i) {
    if (c.kind == DSX_CAT_NONE) return 0;
    return category_raw(c, load_cat_raw(c, i));
}

static inline int check_cat(const dsx_categorizer &c) {
    if (c.kind == DSX_CAT_NONE) return c.ncat == 1 || c.ncat == 0 ? DSX_OK : DSX_ERR_SHAPE;
    if (c.kind < DSX_CAT_CODES || c.kind > DSX_CAT_BINNING) return DSX_ERR_MODE;
    if (!c.col) return DSX_ERR_NULL;
    if (c.ncat < 1) return DSX_ERR_SHAPE;
    if (c.dtype < DSX_F32 || c.dtype > DSX_I64) return DSX_ERR_DTYPE;
    if (c.kind != DSX_CAT_BINNING && (c.dtype == DSX_F32 || c.dtype == DSX_F64))
        return DSX_ERR_DTYPE;
    if (c.kind == DSX_CAT_MODULO && c.modulo == 0) return DSX_ERR_SHAPE;
    if (c.kind == DSX_CAT_BINNING && (c.nbins < 1 || c.binsize == 0.0)) return DSX_ERR_SHAPE;
    return DSX_OK;
}

// Order-preserving unsigned keys of IEEE floats: key(a) < key(b) <=> a < b for non-NaN
// a, b.  -0.0 orders before +0.0.  NaN payloads map to the two extremes, which are the
// "empty" values of the min (all ones) and max (zero) planes; NaN is never inserted.
__device__ __forceinline__ uint32_t fkey32(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unkey32(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}
__device__ __forceinline__ uint64_t fkey64(double d) {
    uint64_t u = (uint64_t)__double_as_longlong(d);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double unkey64(uint64_t k) {
    uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double((long long)u);
}

// ---- exact float64 projection (glyphs/points.py:199-200, line.py:701-702) -----------
// int(x*sx + tx): the product and the sum are rounded separately (no FMA), then the
// result is truncated toward zero.  The bounds test ahead of it keeps the result within
// [-1, W + 1], so a 32-bit conversion is exact.
__device__ __forceinline__ int32_t proj(double x, double s, double t) {
    return (int32_t)__dadd_rn(__dmul_rn(x, s), t);
}

// Axis mapper: identity (LinearAxis) or log10 (LogAxis, core.py:127-145).  log10 is an
// out-of-line call so that the linear path of a kernel's unrolled loops stays small.
static __device__ __noinline__ double axis_log10(double v) { return log10(v); }
__device__ __forceinline__ double axis_map(double v, int32_t log) { return log ? axis_log10(v) : v; }

// ---- aggregate planes ----------------------------------------------------------------
template <int KEYB> struct KeyOf;
template <> struct KeyOf<4> {
    using T = uint32_t;
    static constexpr T kMinEmpty = 0xFFFFFFFFu;
    __device__ static T key(double v) { return fkey32((float)v); }
    __device__ static double val(T k) { return (double)unkey32(k); }
};
template <> struct KeyOf<8> {
    using T = unsigned long long;
    static constexpr T kMinEmpty = ~0ull;
    __device__ static T key(double v) { return (T)fkey64(v); }
    __device__ static double val(T k) { return unkey64((uint64_t)k); }
};

// Which planes a call produces.
enum : uint32_t { PL_CNT = 1, PL_CNTV = 2, PL_SUM = 4, PL_MIN = 8, PL_MAX = 16 };

// The planes a call accumulates: the requested ones, plus sum and count(v) behind a
// finalized mean.
inline uint32_t plane_flags(const dsx_planes *p) {
    return (p->count ? PL_CNT : 0) | (p->countv ? PL_CNTV : 0) | (p->sum ? PL_SUM : 0) |
           (p->min ? PL_MIN : 0) | (p->max ? PL_MAX : 0) | (p->mean ? (PL_SUM | PL_CNTV) : 0);
}

struct Planes {
    uint32_t *count, *countv;
    double *sum;
    void *min, *max;
    double *mean;
};

inline Planes planes_of(const dsx_planes *a) {
    Planes p;
    p.count = a->count; p.countv = a->countv; p.sum = a->sum; p.min = a->min; p.max = a->max;
    p.mean = a->mean;
    return p;
}

// A device pointer that came from memory (a request table, a descriptor array) as a
// global-memory pointer: the compiler cannot prove its address space, and a store through
// a generic pointer is a flat store, which LDS waits after it also wait for.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *global_ptr(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}

// mean's finalize (reductions.py:1294-1297): NaN where the cell has no value
__device__ __forceinline__ double mean_of(double sum, uint32_t nv) {
    return nv > 0u ? sum / (double)nv : (double)NAN;
}

// Global-atomic append of one row (or line pixel) at `cell` (used by the line kernel,
// whose planes the caller initialised).
template <typename TV, bool HAS_VAL>
__device__ __forceinline__ void append(const Planes &p, int64_t cell, TV v) {
    if (p.count) atomicAdd(p.count + cell, 1u);
    if constexpr (HAS_VAL) {
        if (v != v) return;  // NaN: skipped by every value-reading base
        if (p.countv) atomicAdd(p.countv + cell, 1u);
        if (p.sum) unsafeAtomicAdd(p.sum + cell, (double)v);
        using K = KeyOf<sizeof(TV)>;
        if (p.min) atomicMin(reinterpret_cast<typename K::T *>(p.min) + cell, K::key((double)v));
        if (p.max) atomicMax(reinterpret_cast<typename K::T *>(p.max) + cell, K::key((double)v));
    }
}

// Fill n elements of a plane with a kernel (dsx_fill's modes); status as dsx_fill.
int fill_async(void *plane, int32_t mode, int64_t n, hipStream_t stream);

// Number of workgroups for a grid-stride launch over n items.
int grid_for(int64_t n, int block, int per_cu = 8);

}  // namespace dsx
