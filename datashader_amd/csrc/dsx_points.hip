// dsx_points.hip -- point projection + per-bin aggregation (glyphs/points.py:188-221)
// as a tiled LDS engine for gfx950.
//
// Device-scope global atomics run at ~27 G/s on this part while LDS atomics run at
// ~780 G/s (u32 and f64 alike), so no per-row update touches a global atomic:
//
//   select   a sampled coarse density histogram picks the hot rectangle ("window")
//            of the canvas whose aggregate planes fit the 160 KB LDS of a workgroup
//   scan     one streaming pass over the columns: rows that land in the window update
//            the workgroup's LDS copy of the window; every other kept row is appended
//            to its wave's own sub-segment of a spill list of (cell, value) at the
//            wave's running count (a ballot prefix, no atomic, no barrier), and counted
//            in an LDS histogram of canvas tiles (a tile is a run of cells whose planes
//            fit LDS)
//   merge    the per-workgroup window partials are summed cell by cell in a fixed tree
//            order and written to the output planes (plain stores)
//   offsets / scatter
//            the spill segments are counting-sorted by tile; the histograms give every
//            (segment, tile) run its destination, so no atomics are needed
//   tiles    one workgroup per tile accumulates its spill entries in LDS and writes
//            every cell of the tile outside the window (plain stores)
//
// Every output cell is written exactly once; counts, min and max are exact; sums are
// accumulated in float64 (LDS ds_add_f64), so only their last-bit rounding depends on
// the order of the LDS atomics.
#include <math.h>
#include <string.h>

#include "dsx_project.h"

namespace dsx {

constexpr int kScanThreads = 1024;
constexpr int kItems = 4;
constexpr int kChunk = kScanThreads * kItems;  // rows per workgroup iteration
constexpr int kLdsBytes = 160 * 1024;           // LDS of one gfx950 workgroup
constexpr int kCtrlBytes = 64;
constexpr int kFlush = 64;                      // spill entries a wave writes per store
constexpr int kStage = kFlush + 64;             // staged spill entries per wave in LDS
static_assert(kStage >= (kFlush - 1) + 64, "a wave's staging holds a partial line + one item");
constexpr int kPartThreads = 1024;
constexpr int kTileThreads = 512;
// largest LDS of one tile workgroup: half of the CU's LDS less a 2 KB margin, so that two
// tile workgroups always fit on a CU (three when the power-of-two tile takes <= 52 KB).
// Larger tiles mean fewer tiles, so the scatter writes longer runs per tile.
constexpr int kTileLds = 78 * 1024;
constexpr int kMaxTiles = 4096;
constexpr int kMinTileShift = 10;               // smallest tile: 1024 cells
constexpr int kMinPart = 4096;                  // fewest spill entries a tile part is given
constexpr int kScatterPad = 8192;               // slack entries after the spill sub-segments
constexpr int kSelBins = 8192;                  // coarse density histogram bins
constexpr int kSampleRuns = 256;
constexpr int kSampleRun = 4096;

// Internal plane: rows whose value is NaN.  When both count() and count(v) are asked
// for, the kernels keep (count, nan) instead of (count, countv): a row with a value
// then costs one counter update instead of two, and countv = count - nan on output.
constexpr uint32_t PL_NAN = 32;

static uint32_t internal_flags(uint32_t requested) {
    if ((requested & PL_CNT) && (requested & PL_CNTV)) return (requested & ~PL_CNTV) | PL_NAN;
    return requested;
}

// Byte layout of the planes of `cap` cells in one LDS region (or one partial buffer):
// f64 sum first, then keys, then u32 counts, so every plane is naturally aligned.
struct Lay {
    uint32_t flags, cap, o_sum, o_min, o_max, o_cnt, o_cntv, o_nan, bytes;
};

static uint32_t cell_bytes(uint32_t flags, int keyb) {
    return ((flags & PL_SUM) ? 8u : 0u) + ((flags & PL_MIN) ? keyb : 0) +
           ((flags & PL_MAX) ? keyb : 0) + ((flags & PL_CNT) ? 4u : 0u) +
           ((flags & PL_CNTV) ? 4u : 0u) + ((flags & PL_NAN) ? 4u : 0u);
}

static Lay make_lay(uint32_t flags, int keyb, uint32_t cap) {
    Lay L;
    L.flags = flags;
    L.cap = cap;
    uint32_t o = 0;
    L.o_sum = o; o += (flags & PL_SUM) ? 8u * cap : 0u;
    L.o_min = o; o += (flags & PL_MIN) ? (uint32_t)keyb * cap : 0u;
    L.o_max = o; o += (flags & PL_MAX) ? (uint32_t)keyb * cap : 0u;
    L.o_cnt = o; o += (flags & PL_CNT) ? 4u * cap : 0u;
    L.o_cntv = o; o += (flags & PL_CNTV) ? 4u * cap : 0u;
    L.o_nan = o; o += (flags & PL_NAN) ? 4u * cap : 0u;
    L.bytes = (o + 15u) & ~15u;
    return L;
}

// ---- plane regions (LDS or global) --------------------------------------------------------
template <int KEYB>
__device__ __forceinline__ void region_init(uint8_t *r, const Lay &L, int tid, int nt) {
    using K = KeyOf<KEYB>;
    for (uint32_t c = tid; c < L.cap; c += nt) {
        if (L.flags & PL_SUM) reinterpret_cast<double *>(r + L.o_sum)[c] = 0.0;
        if (L.flags & PL_MIN) reinterpret_cast<typename K::T *>(r + L.o_min)[c] = K::kMinEmpty;
        if (L.flags & PL_MAX) reinterpret_cast<typename K::T *>(r + L.o_max)[c] = 0;
        if (L.flags & PL_CNT) reinterpret_cast<uint32_t *>(r + L.o_cnt)[c] = 0u;
        if (L.flags & PL_CNTV) reinterpret_cast<uint32_t *>(r + L.o_cntv)[c] = 0u;
        if (L.flags & PL_NAN) reinterpret_cast<uint32_t *>(r + L.o_nan)[c] = 0u;
    }
}

// LDS atomic update of cell c by one row with value v (vnan: the value is NaN / absent
// for value-reading planes).  `add` is the quantity summed (v, or its squared deviation).
// WF, when not 0, is the region's plane set known at compile time (it must equal L.flags):
// the per-plane tests then fold away.
template <int KEYB, uint32_t WF = 0>
__device__ __forceinline__ void region_add(uint8_t *r, const Lay &L, uint32_t c, bool vnan,
                                           double v, double add) {
    using K = KeyOf<KEYB>;
    const uint32_t f = WF ? WF : L.flags;
    if (f & PL_CNT) atomicAdd(reinterpret_cast<uint32_t *>(r + L.o_cnt) + c, 1u);
    if (vnan) {
        if (f & PL_NAN) atomicAdd(reinterpret_cast<uint32_t *>(r + L.o_nan) + c, 1u);
        return;
    }
    if (f & PL_CNTV) atomicAdd(reinterpret_cast<uint32_t *>(r + L.o_cntv) + c, 1u);
    if (f & PL_SUM) atomicAdd(reinterpret_cast<double *>(r + L.o_sum) + c, add);
    if (f & PL_MIN) atomicMin(reinterpret_cast<typename K::T *>(r + L.o_min) + c, K::key(v));
    if (f & PL_MAX) atomicMax(reinterpret_cast<typename K::T *>(r + L.o_max) + c, K::key(v));
}

// Copy LDS region cell c into the output planes at global cell g.
template <int KEYB>
__device__ __forceinline__ void region_store(const uint8_t *r, const Lay &L, uint32_t c,
                                             const Planes &out, int64_t g) {
    using K = KeyOf<KEYB>;
    if (L.flags & PL_MIN)
        reinterpret_cast<typename K::T *>(out.min)[g] = reinterpret_cast<const typename K::T *>(r + L.o_min)[c];
    if (L.flags & PL_MAX)
        reinterpret_cast<typename K::T *>(out.max)[g] = reinterpret_cast<const typename K::T *>(r + L.o_max)[c];
    uint32_t nv = 0u;  // count(v), when kept
    if (L.flags & PL_CNT) {
        const uint32_t n = reinterpret_cast<const uint32_t *>(r + L.o_cnt)[c];
        if (out.count) out.count[g] = n;
        if (L.flags & PL_NAN) nv = n - reinterpret_cast<const uint32_t *>(r + L.o_nan)[c];
    }
    if (L.flags & PL_CNTV) nv = reinterpret_cast<const uint32_t *>(r + L.o_cntv)[c];
    if (out.countv) out.countv[g] = nv;
    if (L.flags & PL_SUM) {
        const double sum = reinterpret_cast<const double *>(r + L.o_sum)[c];
        if (out.sum) out.sum[g] = sum;
        if (out.mean) out.mean[g] = mean_of(sum, nv);
    }
}

// Add LDS region cell c (a tile with a NaN plane) to the output planes at global cell g,
// which the window merge already wrote: count += cnt, countv += cnt - nan, and the value
// planes combined.  Nothing to do for an empty cell.
template <int KEYB>
__device__ __forceinline__ void region_add_to(const uint8_t *r, const Lay &L, uint32_t c,
                                              const Planes &out, int64_t g) {
    using K = KeyOf<KEYB>;
    using T = typename K::T;
    const uint32_t n = reinterpret_cast<const uint32_t *>(r + L.o_cnt)[c];
    if (n == 0) return;
    const uint32_t nn = reinterpret_cast<const uint32_t *>(r + L.o_nan)[c];
    if (out.count) out.count[g] += n;
    // (the rows a tile holds for a window cell are rows with a NaN value -- rows with a
    // value stay in the window --, so count(v), the sum and a finalized mean are final)
    if (out.countv) out.countv[g] += n - nn;
    if ((L.flags & PL_SUM) && out.sum) out.sum[g] += reinterpret_cast<const double *>(r + L.o_sum)[c];
    if (L.flags & PL_MIN) {
        T &m = reinterpret_cast<T *>(out.min)[g];
        const T x = reinterpret_cast<const T *>(r + L.o_min)[c];
        m = x < m ? x : m;
    }
    if (L.flags & PL_MAX) {
        T &m = reinterpret_cast<T *>(out.max)[g];
        const T x = reinterpret_cast<const T *>(r + L.o_max)[c];
        m = x > m ? x : m;
    }
}

// Fold cell c of another region with layout L (global memory) into region r (LDS); each
// cell is owned by one thread, so no atomics.
template <int KEYB>
__device__ __forceinline__ void region_merge(uint8_t *r, const Lay &L, uint32_t c,
                                             const uint8_t *o) {
    using K = KeyOf<KEYB>;
    using T = typename K::T;
    if (L.flags & PL_SUM)
        reinterpret_cast<double *>(r + L.o_sum)[c] += reinterpret_cast<const double *>(o + L.o_sum)[c];
    if (L.flags & PL_MIN) {
        T &m = reinterpret_cast<T *>(r + L.o_min)[c];
        const T x = reinterpret_cast<const T *>(o + L.o_min)[c];
        m = x < m ? x : m;
    }
    if (L.flags & PL_MAX) {
        T &m = reinterpret_cast<T *>(r + L.o_max)[c];
        const T x = reinterpret_cast<const T *>(o + L.o_max)[c];
        m = x > m ? x : m;
    }
    if (L.flags & PL_CNT)
        reinterpret_cast<uint32_t *>(r + L.o_cnt)[c] += reinterpret_cast<const uint32_t *>(o + L.o_cnt)[c];
    if (L.flags & PL_CNTV)
        reinterpret_cast<uint32_t *>(r + L.o_cntv)[c] += reinterpret_cast<const uint32_t *>(o + L.o_cntv)[c];
    if (L.flags & PL_NAN)
        reinterpret_cast<uint32_t *>(r + L.o_nan)[c] += reinterpret_cast<const uint32_t *>(o + L.o_nan)[c];
}

// ---- window -------------------------------------------------------------------------------
struct Window {
    int32_t x0, y0, w, h;
};

// The window of a call's record {x0, y0, w, h, strategy, ...}, sanitised: anything
// outside the canvas or larger than the LDS capacity (a stale or foreign hint) becomes
// the empty window, and so does any window under the tile strategy.
__device__ __forceinline__ Window load_window(const int32_t *p, int32_t W, int32_t H,
                                              int64_t cap_px) {
    Window w{p[0], p[1], p[2], p[3]};
    if (w.x0 < 0 || w.y0 < 0 || w.w < 0 || w.h < 0 || (int64_t)w.x0 + w.w > W ||
        (int64_t)w.y0 + w.h > H || (int64_t)w.w * w.h > cap_px || p[4] == DSX_STRATEGY_TILES)
        w = Window{0, 0, 0, 0};
    return w;
}

__device__ __forceinline__ bool in_window(const Window &w, int32_t px, int32_t py) {
    return (uint32_t)(px - w.x0) < (uint32_t)w.w && (uint32_t)(py - w.y0) < (uint32_t)w.h;
}

// A call's window record: {x0, y0, w, h, strategy, 0, 0, 0} (kRecWords int32 words).
constexpr int kRecWords = 8;

__global__ void set_window_kernel(int32_t *w, int32_t x0, int32_t y0, int32_t ww, int32_t wh,
                                  int32_t strategy) {
    if (threadIdx.x < kRecWords) {
        const int32_t r[4] = {x0, y0, ww, wh};
        w[threadIdx.x] = threadIdx.x < 4 ? r[threadIdx.x] : (threadIdx.x == 4 ? strategy : 0);
    }
}

// Copy a window record (a call's snapshot of a stored record, or the record a call picked
// back into the caller's store); strategy != 0 overrides the copied strategy.
__global__ void copy_window_kernel(const int32_t *src, int32_t *dst, int32_t strategy) {
    if (threadIdx.x < kRecWords)
        dst[threadIdx.x] = threadIdx.x > 4 ? 0 : (threadIdx.x == 4 && strategy != 0) ? strategy
                                                                               : src[threadIdx.x];
}

// The tile strategy runs (see "the tile strategy" below); any other word: the window one.
__device__ __forceinline__ bool tiles_run(const int32_t *rec) {
    return rec[4] == DSX_STRATEGY_TILES;
}

// Coarse density histogram of a sample of rows: kSampleRuns runs of kSampleRun
// consecutive rows spread evenly over the input; bins of (1 << bshift)^2 pixels.
template <typename TC>
__global__ __launch_bounds__(1024) void select_hist_kernel(PointArgs a, int32_t bshift,
                                                           int32_t cw, int32_t ch,
                                                           uint32_t *hist) {
    __shared__ uint32_t h[kSelBins];
    for (int b = threadIdx.x; b < kSelBins; b += blockDim.x) h[b] = 0u;
    __syncthreads();
    const TC *X = reinterpret_cast<const TC *>(a.x);
    const TC *Y = reinterpret_cast<const TC *>(a.y);
    for (int run = blockIdx.x; run < kSampleRuns; run += gridDim.x) {
        const int64_t start = (int64_t)((double)run * (double)a.n / kSampleRuns);
        for (int k = threadIdx.x; k < kSampleRun; k += blockDim.x) {
            const int64_t i = start + k;
            if (i >= a.n) break;
            int32_t px, py, pc;
            if (!locate<TC>(a, i, X[i], Y[i], px, py, pc)) continue;
            const int32_t bx = px >> bshift, by = py >> bshift;
            if (bx < cw && by < ch) atomicAdd(&h[by * cw + bx], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < cw * ch; b += blockDim.x)
        if (h[b]) atomicAdd(hist + b, h[b]);
}

// Pick the window: for a few aspect ratios, the position (in coarse bins) of the
// rectangle of at most cap_px pixels covering the most sampled rows.
// With tiles_ok, a window holding less than 70% of the sampled rows picks the tile
// strategy instead (no window).
__global__ __launch_bounds__(1024) void select_pick_kernel(const uint32_t *hist, int32_t bshift,
                                                           int32_t cw, int32_t ch, int32_t W,
                                                           int32_t H, int64_t cap_px,
                                                           int32_t tiles_ok, int32_t *win) {
    // (cw+1) x (ch+1) inclusive prefix sums; cw*ch <= kSelBins, so cw + ch <= kSelBins + 1
    __shared__ uint32_t ps[2 * kSelBins + 2];
    __shared__ unsigned long long best[1024 / 64];
    const int pw = cw + 1;
    for (int i = threadIdx.x; i < pw * (ch + 1); i += blockDim.x) ps[i] = 0u;
    __syncthreads();
    for (int i = threadIdx.x; i < cw * ch; i += blockDim.x)
        ps[(i / cw + 1) * pw + (i % cw) + 1] = hist[i];
    __syncthreads();
    for (int y = threadIdx.x + 1; y <= ch; y += blockDim.x)  // rows, in parallel
        for (int x = 1; x <= cw; ++x) ps[y * pw + x] += ps[y * pw + x - 1];
    __syncthreads();
    for (int x = threadIdx.x + 1; x <= cw; x += blockDim.x)  // then columns
        for (int y = 1; y <= ch; ++y) ps[y * pw + x] += ps[(y - 1) * pw + x];
    __syncthreads();
    const int64_t b = 1ll << bshift;
    unsigned long long mybest = 0ull;  // (count << 32) | (0xFFFFFFFF - encoded choice)
    // shapes of about cap_px pixels; a shape's last bin column may overhang the capacity
    // (the window is trimmed to it, pixel by pixel, once chosen)
    constexpr int kAspects = 9;
    const float aspects[kAspects] = {0.25f, 0.35f, 0.5f, 0.7f, 1.0f, 1.4f, 2.0f, 2.8f, 4.0f};
    for (int s = 0; s < kAspects; ++s) {
        int kh = (int)sqrtf((float)cap_px / (aspects[s] * (float)(b * b)));
        if (kh < 1) kh = 1;
        if (kh > ch) kh = ch;
        if ((int64_t)kh * b > cap_px) continue;
        int kw = (int)((cap_px + (int64_t)kh * b * b - 1) / ((int64_t)kh * b * b));
        if (kw > cw) kw = cw;
        const int nx = cw - kw + 1, ny = ch - kh + 1;
        for (int p = threadIdx.x; p < nx * ny; p += blockDim.x) {
            const int x0 = p % nx, y0 = p / nx;
            const uint32_t cnt = ps[(y0 + kh) * pw + (x0 + kw)] - ps[y0 * pw + (x0 + kw)] -
                                 ps[(y0 + kh) * pw + x0] + ps[y0 * pw + x0];
            const uint32_t code = ((uint32_t)s << 28) | ((uint32_t)y0 << 14) | (uint32_t)x0;
            const unsigned long long key = ((unsigned long long)cnt << 32) | (0xFFFFFFFFu - code);
            if (key > mybest) mybest = key;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(mybest, off, 64);
        mybest = o > mybest ? o : mybest;
    }
    if ((threadIdx.x & 63) == 0) best[threadIdx.x >> 6] = mybest;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long k = 0ull;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) k = best[w] > k ? best[w] : k;
        for (int i = 4; i < kRecWords; ++i) win[i] = 0;
        win[4] = DSX_STRATEGY_WINDOW;
        const uint32_t sampled = ps[ch * pw + cw];
        if (tiles_ok && (k >> 32) * 10ull < (unsigned long long)sampled * 7ull) {
            win[0] = 0; win[1] = 0; win[2] = 0; win[3] = 0;
            win[4] = DSX_STRATEGY_TILES;
            return;
        }
        if (k == 0ull) {  // no shape fits a coarse bin: no window, every row spills
            win[0] = 0; win[1] = 0; win[2] = 0; win[3] = 0;
            return;
        }
        const uint32_t code = 0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFu);
        const int s = (int)(code >> 28), y0 = (int)((code >> 14) & 0x3FFF), x0 = (int)(code & 0x3FFF);
        int kh = (int)sqrtf((float)cap_px / (aspects[s] * (float)(b * b)));
        if (kh < 1) kh = 1;
        if (kh > ch) kh = ch;
        int kw = (int)((cap_px + (int64_t)kh * b * b - 1) / ((int64_t)kh * b * b));
        if (kw > cw) kw = cw;
        if (kw < 1) kw = 1;
        int px0 = (int)(x0 * b);
        const int py0 = (int)(y0 * b);
        int ww = (int)(kw * b), wh = (int)(kh * b);
        if (px0 + ww > W) ww = W - px0;
        if (py0 + wh > H) wh = H - py0;
        if (wh >= 1 && (int64_t)ww * wh > cap_px) {  // trim the overhang, about the centre
            const int tw = (int)(cap_px / wh);
            px0 += (ww - tw) / 2;
            ww = tw;
        }
        if ((int64_t)ww * wh > cap_px || ww < 1 || wh < 1) { ww = 0; wh = 0; }
        // the coarse shape leaves up to a bin row / column of capacity unused: widen,
        // then heighten, the window pixel by pixel around its centre to use it
        int wx0 = px0, wy0 = py0;
        if (ww > 0) {
            int64_t mw = cap_px / wh;
            mw = mw > W ? W : mw;
            if (mw > ww) {
                wx0 -= (int)((mw - ww) / 2);
                ww = (int)mw;
                wx0 = wx0 < 0 ? 0 : (wx0 + ww > W ? W - ww : wx0);
            }
            int64_t mh = cap_px / ww;
            mh = mh > H ? H : mh;
            if (mh > wh) {
                wy0 -= (int)((mh - wh) / 2);
                wh = (int)mh;
                wy0 = wy0 < 0 ? 0 : (wy0 + wh > H ? H - wh : wy0);
            }
        }
        win[0] = wx0; win[1] = wy0; win[2] = ww; win[3] = wh;
    }
}

// ---- scan ------------------------------------------------------------------------------------
// Workgroup g of the scan handles chunks g, g + G, ...; wave w of it appends the rows it
// spills to its own sub-segment [(g*kWaves + w)*seg_cap, ... + seg_count[g*kWaves + w])
// of the spill list, so no reservation is shared by two waves (the running count stays
// in a scalar register).
constexpr int kWaves = kScanThreads / 64;

struct ScanArgs {
    Lay wl;                          // window planes (cap = window capacity in cells)
    const int32_t *window;           // device [x0, y0, w, h]
    uint8_t *partials;               // gridDim.x regions of wl.bytes
    uint32_t *spill_cell;            // spill sub-segments
    void *spill_val;
    uint32_t *seg_count;             // [G * kWaves] entries in each sub-segment
    int64_t seg_cap;                 // entries a sub-segment can hold (its wave's rows)
    uint32_t *hist;                  // [G][ntiles] spilled entries per tile
    uint32_t tile_cells, tile_shift;  // tile_cells = 1 << tile_shift
    int32_t ntiles;
    uint32_t o_thist, o_stage_cell, o_stage_val;  // LDS byte offsets
    uint32_t spill_nan;              // rows with a NaN value are spilled (count with countv)
    const double *mean;              // squared-deviation mode (m2 pass)
};

// How the scan reads a row's category: no categorizer, a one-byte code column (int8 /
// uint8), or any other categorizer column.
// CK_CODE8: a one-byte code column read by category codes (no modulo / binning).
enum : int { CK_NONE = 0, CK_BYTE = 1, CK_ANY = 2, CK_CODE8 = 3 };

// Rows of chunk ck handled by this thread.  Every load is issued unconditionally (a row
// past the end reads the last row and is dropped when binned), so the loads of a chunk
// are a fixed-length run and the compiler waits for exactly the chunk it bins while the
// chunks prefetched behind it stay in flight.  Loaded values are not touched here.
// WIDE (float32 columns, 16-byte aligned, n a multiple of 4 -- checked on the host): the
// thread's rows are b + 4 tid .. b + 4 tid + 3, one 16-byte load per column (the streams
// then run ~10 % faster than with a 4-byte load per row: 6.95 against 6.24 TB/s for three
// columns on one MI355X).
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename TC, typename TV, bool HAS_VAL, int CK, bool WIDE = false>
struct Rows {
    TC x[kItems], y[kItems];
    TV v[kItems];
    int64_t c[kItems];

    __device__ __forceinline__ void load(const PointArgs &a, int64_t ck, int tid) {
        if constexpr (WIDE) {
            static_assert(sizeof(TC) == 4 && (!HAS_VAL || sizeof(TV) == 4) &&
                          (CK == CK_NONE || CK == CK_CODE8) && kItems == 4,
                          "wide rows: float32 columns (one-byte codes), four rows a thread");
            int64_t b = ck * kChunk;
            uint32_t lastq;  // the last readable group of four rows
            if (b >= a.n) {  // a prefetch past the end: re-read the last four rows
                b = a.n - 4;
                lastq = 0;
            } else {
                const int64_t rem = a.n - b;
                lastq = rem >= kChunk ? (uint32_t)(kChunk / 4 - 1) : (uint32_t)(rem / 4 - 1);
            }
            const uint32_t q = (uint32_t)tid < lastq ? (uint32_t)tid : lastq;
            const f32x4 X = stream_load(reinterpret_cast<const f32x4 *>(reinterpret_cast<const float *>(a.x) + b) + q);
            const f32x4 Y = stream_load(reinterpret_cast<const f32x4 *>(reinterpret_cast<const float *>(a.y) + b) + q);
            f32x4 V = {0.f, 0.f, 0.f, 0.f};
            if constexpr (HAS_VAL)
                V = stream_load(reinterpret_cast<const f32x4 *>(reinterpret_cast<const float *>(a.val) + b) + q);
            uint32_t C4 = 0u;  // the four rows' one-byte codes (a 4-byte aligned column)
            if constexpr (CK == CK_CODE8)
                C4 = reinterpret_cast<const uint32_t *>(reinterpret_cast<const uint8_t *>(a.cat.col) + b)[q];
#pragma unroll
            for (int it = 0; it < kItems; ++it) {
                x[it] = X[it];
                y[it] = Y[it];
                v[it] = HAS_VAL ? (TV)V[it] : TV(0);
                c[it] = CK == CK_CODE8 ? (int64_t)((C4 >> (8 * it)) & 0xFFu) : 0;
            }
            return;
        }
        // chunk base and last readable in-chunk row are uniform; each thread adds a
        // 32-bit offset (scalar base + vector offset addressing)
        int64_t b = ck * kChunk;
        uint32_t lastj;
        if (b >= a.n) {  // a prefetch past the end: re-read the last row
            b = a.n > 0 ? a.n - 1 : 0;
            lastj = 0;
        } else {
            const int64_t rem = a.n - b;
            lastj = rem >= kChunk ? (uint32_t)(kChunk - 1) : (uint32_t)(rem - 1);
        }
        const TC *X = reinterpret_cast<const TC *>(a.x) + b;
        const TC *Y = reinterpret_cast<const TC *>(a.y) + b;
        const TV *V = reinterpret_cast<const TV *>(a.val) + (HAS_VAL ? b : 0);
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            uint32_t j = (uint32_t)(it * kScanThreads + tid);
            j = j < lastj ? j : lastj;
            x[it] = stream_load(X + j);
            y[it] = stream_load(Y + j);
            if constexpr (HAS_VAL) v[it] = stream_load(V + j);
            else v[it] = TV(0);
            // (a byte column's 64-byte wave reads share cache lines between waves: cached)
            if constexpr (CK == CK_BYTE || CK == CK_CODE8)
                c[it] = reinterpret_cast<const uint8_t *>(a.cat.col)[b + j];
            else if constexpr (CK == CK_ANY) c[it] = load_cat_raw(a.cat, b + j);
            else c[it] = 0;
        }
    }
};

// float64 to int32 by the hardware conversion (truncation; out-of-range values saturate,
// NaN gives 0): defined for every input, so a row outside the view needs no select before
// it (its cell is dropped).  (A C++ conversion of an out-of-range value is undefined.)
__device__ __forceinline__ int32_t cvt_sat_i32(double d) {
    int32_t r;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(d));
    return r;
}

// a * b + c in 32 bits, the multiply kept apart from the add: fused, they become a 64-bit
// multiply-add whose unused high addend may share a register with a load still in flight,
// and the wait for that load drains the scan's prefetch.
__device__ __forceinline__ uint32_t mul_add32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t m = a * b;
    asm volatile("" : "+v"(m));
    return m + c;
}

// Category of a scanned row from its loaded raw input; -1 when dropped.
template <int CK>
__device__ __forceinline__ int32_t scan_category(const Cat &c, int64_t raw) {
    if constexpr (CK == CK_NONE) return 0;
    if constexpr (CK == CK_CODE8) {  // category_raw's codes branch, in 32 bits
        int32_t k = c.dtype == DSX_I8 ? (int32_t)(int8_t)(uint8_t)raw : (int32_t)raw;
        if (k < 0) k += c.ncat;
        return (uint32_t)k < (uint32_t)c.ncat ? k : -1;
    }
    if constexpr (CK == CK_BYTE) raw = c.dtype == DSX_I8 ? (int64_t)(int8_t)(uint8_t)raw : raw;
    return (int32_t)category_raw(c, raw);
}

// LIN: linear axes (no log10 call in the loop); WF: the window's plane set when known at
// compile time (0: read from the layout).
template <typename TC, typename TV, bool HAS_VAL, int KEYB, bool SQDEV, int CK, bool LIN = false,
          uint32_t WF = 0, bool WIDE = false>
__global__ __launch_bounds__(kScanThreads) void scan_kernel(PointArgs a, ScanArgs s) {
    if (tiles_run(s.window)) return;  // uniform: the tile strategy's scan runs instead
    extern __shared__ __align__(16) uint8_t smem[];
    uint8_t *win = smem;
    uint32_t *thist = reinterpret_cast<uint32_t *>(smem + s.o_thist);  // [ntiles]
    const int tid = threadIdx.x;
    const int ncat = a.cat.ncat;
    const Window w = load_window(s.window, a.W, a.H, (int64_t)(s.wl.cap / (uint32_t)ncat));
    region_init<KEYB>(win, s.wl, tid, kScanThreads);
    for (int t = tid; t < s.ntiles; t += kScanThreads) thist[t] = 0u;
    __syncthreads();
    const int wave = tid >> 6;
    const size_t seg0 = ((size_t)blockIdx.x * kWaves + wave) * (size_t)s.seg_cap;
    // spills are staged per wave in LDS and written kFlush at a time as whole cache lines
    uint32_t *st_cell = reinterpret_cast<uint32_t *>(smem + s.o_stage_cell) + wave * kStage;
    TV *st_val = reinterpret_cast<TV *>(smem + s.o_stage_val) + wave * kStage;
    uint32_t staged = 0, flushed = 0;  // wave-uniform
    TV *spill_val = reinterpret_cast<TV *>(s.spill_val);
    const int lane = threadIdx.x & 63;
    const int64_t nchunks = (a.n + kChunk - 1) / kChunk;
    // software pipeline, two chunks deep: while chunk ck is binned, the loads of chunks
    // ck + G and ck + 2G are in flight.  Three register sets rotate roles (binned / in
    // flight / being loaded), so no set is ever copied: a copy would wait for its loads.
    using R = Rows<TC, TV, HAS_VAL, CK, WIDE>;
    R ra, rb, rc;
    ra.load(a, blockIdx.x, tid);
    rb.load(a, (int64_t)blockIdx.x + gridDim.x, tid);
    int64_t ck = blockIdx.x;
    auto step = [&](R &cur, R &far) {
        far.load(a, ck + 2 * (int64_t)gridDim.x, tid);
        const int64_t rem64 = a.n - ck * kChunk;  // rows of this chunk, uniform
        const uint32_t rem = rem64 >= kChunk ? (uint32_t)kChunk
                                             : (rem64 > 0 ? (uint32_t)rem64 : 0u);
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            // the row's place in the chunk (rows past rem are dropped)
            const uint32_t j = WIDE ? (uint32_t)(4 * tid + it) : (uint32_t)(it * kScanThreads + tid);
            int32_t px = 0, py = 0, pc = 0;
            bool keep;
            if constexpr (LIN) {
                // locate_xy without branches, for a view whose translation is exactly
                // -xmin*sx, -ymin*sy (checked on the host): an in-range row then projects
                // into [0, W] -- x >= xmin gives fl(x*sx) >= fl(xmin*sx) -- so the
                // wraparound and the lower bound never apply; a row out of range (NaN
                // included) converts a harmless 0 and is dropped by `inr`
                const double x = (double)cur.x[it], y = (double)cur.y[it];
                const bool inr = (a.xmin <= x) & (x <= a.xmax) & (a.ymin <= y) & (y <= a.ymax);
                const double fx = inr ? __dadd_rn(__dmul_rn(x, a.sx), a.tx) : 0.0;
                const double fy = inr ? __dadd_rn(__dmul_rn(y, a.sy), a.ty) : 0.0;
                const int32_t xx = (int32_t)fx, yy = (int32_t)fy;
                px = xx < a.W - 1 ? xx : a.W - 1;
                py = yy < a.H - 1 ? yy : a.H - 1;
                keep = (j < rem) & inr;
            } else {
                keep = j < rem && locate_xy<TC>(a, cur.x[it], cur.y[it], px, py);
            }
            if constexpr (CK != CK_NONE) {
                pc = scan_category<CK>(a.cat, cur.c[it]);
                keep = keep && pc >= 0;
            }
            const double v = HAS_VAL ? (double)cur.v[it] : 0.0;
            const bool vnan = HAS_VAL ? (v != v) : true;
            const bool inw = keep && in_window(w, px, py) && !(HAS_VAL && s.spill_nan && vnan);
            if (inw) {
                // 24-bit multiplies (full rate): window and canvas sides are < 2^24
                const uint32_t lc = CK == CK_NONE
                    ? __umul24((uint32_t)(py - w.y0), (uint32_t)w.w) + (uint32_t)(px - w.x0)
                    : (uint32_t)(((py - w.y0) * w.w + (px - w.x0)) * ncat + pc);
                double add = v;
                if constexpr (SQDEV) {
                    const int64_t gc = ((int64_t)py * a.W + px) * ncat + pc;
                    const double d = v - s.mean[gc];
                    add = d * d;
                }
                region_add<KEYB, WF>(win, s.wl, lc, vnan, v, add);
            }
            // spill: staged at the wave's running count, then appended to its sub-segment
            const bool spill = keep && !inw;
            const unsigned long long m = __ballot(spill);
            if (spill) {
                // cells < 2^32 (checked on the host), so 32-bit arithmetic is exact
                const uint32_t cell = CK == CK_NONE
                    ? __umul24((uint32_t)py, (uint32_t)a.W) + (uint32_t)px
                    : ((uint32_t)py * (uint32_t)a.W + (uint32_t)px) * (uint32_t)ncat + (uint32_t)pc;
                const uint32_t slot = staged + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                st_cell[slot] = cell;
                if constexpr (HAS_VAL) st_val[slot] = cur.v[it];
            }
            staged += (uint32_t)__popcll(m);
            // staged < kFlush before the item, so staged < kFlush + 64 <= kStage here;
            // write whole lines until fewer than kFlush entries remain
            while (staged >= (uint32_t)kFlush) {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                if (lane < kFlush) {
                    const size_t slot = (size_t)flushed + lane;
                    const uint32_t c = st_cell[lane];
                    TV vv = TV(0);
                    if constexpr (HAS_VAL) vv = st_val[lane];
                    if (slot < (size_t)s.seg_cap) {
                        s.spill_cell[seg0 + slot] = c;
                        if constexpr (HAS_VAL) spill_val[seg0 + slot] = vv;
                        // counted here, with every lane busy, rather than per binned row
                        atomicAdd(&thist[c >> s.tile_shift], 1u);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                const uint32_t rest = staged - kFlush;  // < 64: move it to the front
                if ((uint32_t)lane < rest) {
                    st_cell[lane] = st_cell[kFlush + lane];
                    if constexpr (HAS_VAL) st_val[lane] = st_val[kFlush + lane];
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                flushed += kFlush;
                staged = rest;
            }
        }
        ck += gridDim.x;
    };
    // no exit between the three steps: a loop edge that skips steps makes the compiler's
    // wait counts assume fewer loads in flight, and it then drains the prefetch once per
    // iteration.  Steps past the last chunk load the last row and keep nothing (rem = 0).
    while (ck < nchunks) {
        step(ra, rc);
        step(rb, ra);
        step(rc, rb);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if ((uint32_t)lane < staged) {  // the last partial line
        const size_t slot = (size_t)flushed + lane;
        if (slot < (size_t)s.seg_cap) {
            const uint32_t c = st_cell[lane];
            s.spill_cell[seg0 + slot] = c;
            if constexpr (HAS_VAL) spill_val[seg0 + slot] = st_val[lane];
            atomicAdd(&thist[c >> s.tile_shift], 1u);
        }
    }
    __syncthreads();
    // this workgroup's window partial, tile histogram and segment length
    uint4 *dst = reinterpret_cast<uint4 *>(s.partials + (size_t)blockIdx.x * s.wl.bytes);
    const uint4 *src = reinterpret_cast<const uint4 *>(win);
    for (uint32_t j = tid; j < s.wl.bytes / 16u; j += kScanThreads) dst[j] = src[j];
    for (int t = tid; t < s.ntiles; t += kScanThreads)
        s.hist[(size_t)blockIdx.x * s.ntiles + t] = thist[t];
    if (lane == 0) {
        const uint32_t filled = flushed + staged;
        s.seg_count[(size_t)blockIdx.x * kWaves + wave] =
            filled < (uint32_t)s.seg_cap ? filled : (uint32_t)s.seg_cap;
    }
}

// ---- merge the window partials --------------------------------------------------------------
// A block covers kMergeCells cells; its kMergeWays thread rows each sum a strided subset
// of the partials, then the rows are combined in a fixed tree order (deterministic).
constexpr int kMergeCells = 64;
constexpr int kMergeWays = 16;

// WF: the plane set when known at compile time (0: read from the layout); known, the
// loop's loads carry no branches and four partials' loads are in flight at a time.
template <int KEYB, uint32_t WF = 0>
__global__ __launch_bounds__(kMergeCells * kMergeWays) void merge_window_kernel(
    Lay L, const uint8_t *partials, int nparts, const int32_t *window, int32_t W, int32_t H,
    int32_t ncat, Planes out) {
    using K = KeyOf<KEYB>;
    __shared__ double s_sum[kMergeWays][kMergeCells];
    __shared__ typename K::T s_min[kMergeWays][kMergeCells];
    __shared__ typename K::T s_max[kMergeWays][kMergeCells];
    __shared__ uint32_t s_cnt[kMergeWays][kMergeCells];
    __shared__ uint32_t s_cntv[kMergeWays][kMergeCells];
    __shared__ uint32_t s_nan[kMergeWays][kMergeCells];
    const uint32_t f = WF ? WF : L.flags;
    const Window w = load_window(window, W, H, (int64_t)(L.cap / (uint32_t)ncat));
    const uint32_t ncells = (uint32_t)(w.w * w.h * ncat);
    const int cl = threadIdx.x % kMergeCells, way = threadIdx.x / kMergeCells;
    const uint32_t c = blockIdx.x * kMergeCells + cl;
    const bool live = c < ncells && c < L.cap;
    double sum = 0.0;
    typename K::T mn = K::kMinEmpty, mx = 0;
    uint32_t cnt = 0u, cntv = 0u, nan = 0u;
    if (live) {
#pragma unroll 4
        for (int g = way; g < nparts; g += kMergeWays) {
            const uint8_t *r = partials + (size_t)g * L.bytes;
            if (f & PL_SUM) sum += reinterpret_cast<const double *>(r + L.o_sum)[c];
            if (f & PL_MIN) {
                const typename K::T k = reinterpret_cast<const typename K::T *>(r + L.o_min)[c];
                mn = k < mn ? k : mn;
            }
            if (f & PL_MAX) {
                const typename K::T k = reinterpret_cast<const typename K::T *>(r + L.o_max)[c];
                mx = k > mx ? k : mx;
            }
            if (f & PL_CNT) cnt += reinterpret_cast<const uint32_t *>(r + L.o_cnt)[c];
            if (f & PL_CNTV) cntv += reinterpret_cast<const uint32_t *>(r + L.o_cntv)[c];
            if (f & PL_NAN) nan += reinterpret_cast<const uint32_t *>(r + L.o_nan)[c];
        }
    }
    s_sum[way][cl] = sum; s_min[way][cl] = mn; s_max[way][cl] = mx;
    s_cnt[way][cl] = cnt; s_cntv[way][cl] = cntv; s_nan[way][cl] = nan;
    __syncthreads();
    for (int half = kMergeWays / 2; half > 0; half >>= 1) {
        if (way < half) {
            s_sum[way][cl] += s_sum[way + half][cl];
            s_min[way][cl] = s_min[way + half][cl] < s_min[way][cl] ? s_min[way + half][cl] : s_min[way][cl];
            s_max[way][cl] = s_max[way + half][cl] > s_max[way][cl] ? s_max[way + half][cl] : s_max[way][cl];
            s_cnt[way][cl] += s_cnt[way + half][cl];
            s_cntv[way][cl] += s_cntv[way + half][cl];
            s_nan[way][cl] += s_nan[way + half][cl];
        }
        __syncthreads();
    }
    if (way != 0 || !live) return;
    const uint32_t pix = c / (uint32_t)ncat, cat = c - pix * (uint32_t)ncat;
    const int32_t ly = (int32_t)(pix / (uint32_t)w.w), lx = (int32_t)(pix - (uint32_t)ly * w.w);
    const int64_t g = ((int64_t)(w.y0 + ly) * W + (w.x0 + lx)) * ncat + cat;
    if (f & PL_MIN) reinterpret_cast<typename K::T *>(out.min)[g] = s_min[0][cl];
    if (f & PL_MAX) reinterpret_cast<typename K::T *>(out.max)[g] = s_max[0][cl];
    uint32_t nv = 0u;
    if (f & PL_CNT) {
        if (out.count) out.count[g] = s_cnt[0][cl];
        // count() with count(v): the window holds the rows with a value only (rows with
        // a NaN value are spilled), so its count is also its count(v); the tile adds the
        // spilled rows of the cell afterwards
        nv = (f & PL_NAN) ? s_cnt[0][cl] - s_nan[0][cl] : s_cnt[0][cl];
    }
    if (f & PL_CNTV) nv = s_cntv[0][cl];
    if (out.countv) out.countv[g] = nv;
    if (f & PL_SUM) {
        if (out.sum) out.sum[g] = s_sum[0][cl];
        if (out.mean) out.mean[g] = mean_of(s_sum[0][cl], nv);
    }
}

// ---- partition the spill list by tile --------------------------------------------------------
// The spill list arrives as G per-workgroup segments with their tile histograms
// (written by the scan); group g of the partition owns segment g.
struct PartArgs {
    const uint32_t *cell;
    const void *val;
    const uint32_t *seg_count;        // [G * kWaves]
    int64_t seg_cap;
    int32_t ngroups;                  // G (<= kPartThreads)
    uint32_t tile_cells, tile_shift;  // tile_cells = 1 << tile_shift
    int32_t ntiles;
    uint32_t *hist;                   // [G][ntiles]
    unsigned long long *tile_start;   // [ntiles + 1]
    uint32_t *sorted_cell;            // [sink + 64 * G]: the sorted spill, then one
    void *sorted_val;                 // 64-slot sink line per scatter workgroup
    size_t sink;
    // tile work split: tile t is processed by parts [part_start[t], part_start[t + 1]),
    // over which tile t's entries are split evenly (*part_len: the largest part size)
    uint32_t *part_start;             // [ntiles + 1]
    uint32_t *part_tile;              // [max_parts] tile of each part
    int32_t max_parts;
    uint32_t *part_len;               // [1]
    uint8_t *scratch;                 // one tile-layout region per part of a split tile
    const int32_t *rec;               // the call's window record (its strategy)
    uint32_t min_part_tiles;          // fewest blocks a tile part is given (tile strategy)
};

// Block-wide exclusive scan of one value per thread (1024 threads).
__device__ __forceinline__ unsigned long long block_exscan(unsigned long long v,
                                                           unsigned long long *sh,
                                                           unsigned long long &total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            const unsigned long long t = sh[k];
            sh[k] = run;
            run += t;
        }
        sh[blockDim.x >> 6] = run;
    }
    __syncthreads();
    total = sh[blockDim.x >> 6];
    const unsigned long long r = sh[wv] + x - v;
    __syncthreads();
    return r;
}

// Block-wide exclusive scan of one 32-bit value per thread (1024 threads, totals below
// 2^32): a wave scan, then wave 0 scans the 16 wave totals (no serial loop).
__device__ __forceinline__ uint32_t block_exscan32(uint32_t v, uint32_t *sh, uint32_t &total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    if (wv == 0) {
        constexpr int kW = kPartThreads / 64;
        const uint32_t w = lane < kW ? sh[lane] : 0u;
        uint32_t z = w;
#pragma unroll
        for (int off = 1; off < kW; off <<= 1) {
            const uint32_t y = __shfl_up(z, off, 64);
            if (lane >= off) z += y;
        }
        if (lane < kW) sh[lane] = z - w;            // exclusive prefix of the wave totals
        if (lane == kW - 1) sh[kW] = z;             // block total
    }
    __syncthreads();
    total = sh[kPartThreads / 64];
    const uint32_t r = sh[wv] + x - v;
    __syncthreads();
    return r;
}

// Per tile (one workgroup each): exclusive prefix of the partition groups' counts, in
// place, and the tile total.
__global__ __launch_bounds__(kPartThreads) void spill_colscan_kernel(PartArgs p,
                                                                    unsigned long long *totals) {
    __shared__ unsigned long long sh[kPartThreads / 64 + 1];
    const int t = blockIdx.x;
    const int g = threadIdx.x;  // ngroups <= kPartThreads
    const unsigned long long v = g < p.ngroups ? p.hist[(size_t)g * p.ntiles + t] : 0ull;
    unsigned long long tot;
    const unsigned long long ex = block_exscan(v, sh, tot);
    if (g < p.ngroups) p.hist[(size_t)g * p.ntiles + t] = (uint32_t)ex;
    if (g == 0) totals[t] = tot;
}

__global__ __launch_bounds__(kPartThreads) void spill_tilescan_kernel(PartArgs p,
                                                                     const unsigned long long *totals) {
    __shared__ unsigned long long sh[kPartThreads / 64 + 1];
    unsigned long long carry = 0;
    for (int t0 = 0; t0 < p.ntiles; t0 += blockDim.x) {
        const int t = t0 + threadIdx.x;
        const unsigned long long v = t < p.ntiles ? totals[t] : 0ull;
        unsigned long long tot;
        const unsigned long long ex = block_exscan(v, sh, tot);
        if (t < p.ntiles) p.tile_start[t] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) p.tile_start[p.ntiles] = carry;
    // parts of at most E = max(kMinPart, 1.5 * mean) entries (a tile's entries are split
    // evenly over its parts): at most ntiles + 2 * ntiles / 3 parts in all, and a tile
    // heavier than the mean (the edge of a dense cluster just outside the window) is
    // spread over several workgroups
    // (under the tile strategy the counts are blocks of entries, not entries)
    const unsigned long long minp = tiles_run(p.rec) ? p.min_part_tiles : (unsigned long long)kMinPart;
    const unsigned long long mean = (carry + p.ntiles - 1) / p.ntiles;
    const unsigned long long e15 = mean + (mean + 1) / 2;
    const unsigned long long E = e15 > minp ? e15 : minp;
    if (threadIdx.x == 0) *p.part_len = (uint32_t)(E < 0xFFFFFFFFull ? E : 0xFFFFFFFFull);
    unsigned long long pcarry = 0;
    for (int t0 = 0; t0 < p.ntiles; t0 += blockDim.x) {
        const int t = t0 + threadIdx.x;
        unsigned long long np = 0;
        if (t < p.ntiles) {
            np = (totals[t] + E - 1) / E;
            np = np < 1 ? 1 : np;
        }
        unsigned long long tot;
        const unsigned long long ex = block_exscan(np, sh, tot);
        if (t < p.ntiles) {
            p.part_start[t] = (uint32_t)(pcarry + ex);
            for (unsigned long long q = 0; q < np; ++q)
                if (pcarry + ex + q < (unsigned long long)p.max_parts)  // always (bound above)
                    p.part_tile[pcarry + ex + q] = (uint32_t)t;
        }
        pcarry += tot;
    }
    if (threadIdx.x == 0) p.part_start[p.ntiles] = (uint32_t)pcarry;
}

// Scatter: workgroup g sorts the spill of scan workgroup g by tile.  Wave w reads
// sub-segment w (written by scan wave w), kWv entries a round; each round is counting-
// sorted by tile in LDS (LDS counters rank the entries, one block scan over the tiles
// places them) and written out in sorted order, so a tile's entries of the round leave
// as one run.  All rounds are alike: every load is issued (past the end of a sub-segment
// it re-reads the last entry) and every thread writes exactly kPer entries (slots past
// the round's count go to the workgroup's own sink line after the sorted spill), so the
// compiler counts the loads and stores in flight exactly, and the next round's loads (two
// register sets alternate) stay in flight behind this round's stores.
template <typename TV, bool HAS_VAL>
struct ScatterCfg {  // entries per round: cell + value staging fits LDS next to the tile arrays,
                     // and two register sets of them fit 128 VGPRs (no spills)
    static constexpr int kSub = !HAS_VAL ? 20480 : (sizeof(TV) == 8 ? 8192 : 12288);
    static constexpr int kPer = kSub / kPartThreads;  // per thread
    static constexpr int kWv = kSub / kWaves;         // per wave
};
constexpr int kTpt = kMaxTiles / kPartThreads;         // tiles per thread in the block scan

template <typename TV, bool HAS_VAL>
struct ScatterRound {
    uint32_t c[ScatterCfg<TV, HAS_VAL>::kPer];
    TV v[ScatterCfg<TV, HAS_VAL>::kPer];
};

template <typename TV, bool HAS_VAL>
__global__ __launch_bounds__(kPartThreads) void spill_scatter_kernel(PartArgs p) {
    using C = ScatterCfg<TV, HAS_VAL>;
    constexpr int kSub = C::kSub, kPer = C::kPer, kWv = C::kWv;
    static_assert(kPer * kPartThreads == kSub && kWv % 64 == 0, "round shape");
    static_assert(kTpt * kPartThreads == kMaxTiles, "tile split");
    static_assert(3 * kWv <= kScatterPad, "loads past a sub-segment stay in the slack");
    if (tiles_run(p.rec)) return;  // uniform: no spill list under the tile strategy
    __shared__ uint32_t cnt[kMaxTiles];  // round counts, then the round start of each tile
    __shared__ uint32_t wb[kMaxTiles];   // sorted position of round slot 0, per tile
    __shared__ uint32_t s_cell[kSub];
    __shared__ TV s_val[HAS_VAL ? kSub : 1];
    __shared__ uint32_t sh[kPartThreads / 64 + 1];
    __shared__ uint32_t s_rounds;
    const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar bases
    // tiles [tid * kTpt, tid * kTpt + kTpt): where group g's next entry of each goes in
    // the sorted spill (positions fit 32 bits: the host rejects 2^32 rows or more)
    uint32_t cur[kTpt];
#pragma unroll
    for (int i = 0; i < kTpt; ++i) {
        const int t = tid * kTpt + i;
        cur[i] = t < p.ntiles ? (uint32_t)(p.tile_start[t] + p.hist[(size_t)g * p.ntiles + t]) : 0u;
        cnt[t] = 0u;
    }
    uint32_t len = p.seg_count[(size_t)g * kWaves + wave];  // my wave's sub-segment
    if ((int64_t)len > p.seg_cap) len = (uint32_t)p.seg_cap;
    if (tid == 0) s_rounds = 0u;
    __syncthreads();
    if (lane == 0) atomicMax(&s_rounds, (len + kWv - 1) / (uint32_t)kWv);
    __syncthreads();
    const uint32_t rounds = s_rounds;
    if (rounds == 0) return;  // uniform: this group spilled nothing
    // Loads run past the end of a sub-segment into the next one, and past the last one
    // into kScatterPad slack entries (never used: masked by len).  Sub-segments are
    // shorter than 2^31 entries (rows per wave), so int32 round arithmetic is exact.
    const size_t seg = ((size_t)g * kWaves + wave) * (size_t)p.seg_cap;
    const uint32_t *cellp = p.cell + seg;
    const TV *valp = reinterpret_cast<const TV *>(p.val) + (HAS_VAL ? seg : 0);
    TV *sval = reinterpret_cast<TV *>(p.sorted_val);
    const size_t sink = p.sink + (size_t)g * 64 + lane;
    using R = ScatterRound<TV, HAS_VAL>;
    auto load = [&](R &x, int32_t r) {  // one vector address, immediate offsets
        // a round wholly past my sub-segment (the prefetch after the last round, an
        // empty round) re-reads the first kWv entries of the list, which stay cached
        const bool live = r * kWv < (int32_t)len;
        const uint32_t *q = live ? cellp + ((size_t)r * kWv + lane) : p.cell + lane;
        const TV *qv = live ? valp + (HAS_VAL ? (size_t)r * kWv + lane : 0)
                            : reinterpret_cast<const TV *>(p.val) + (HAS_VAL ? lane : 0);
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            x.c[k] = q[k * 64];
            if constexpr (HAS_VAL) x.v[k] = qv[k * 64];
            else x.v[k] = TV(0);
        }
    };
    auto round = [&](const R &x, int32_t r) {
        const int32_t lim = (int32_t)len - r * kWv;  // entries of this round left to my wave
        uint32_t rr[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {  // rank within the tile
            rr[k] = 0xFFFFFFFFu;
            const uint32_t t = x.c[k] >> p.tile_shift;
            if (lane < lim - k * 64 && t < (uint32_t)p.ntiles)  // always in range, for entries the scan wrote
                rr[k] = atomicAdd(&cnt[t], 1u);
        }
        __syncthreads();
        uint32_t v[kTpt], sum = 0u;
#pragma unroll
        for (int i = 0; i < kTpt; ++i) {
            v[i] = cnt[tid * kTpt + i];  // 0 past the last tile
            sum += v[i];
        }
        uint32_t total;
        uint32_t run = block_exscan32(sum, sh, total);
#pragma unroll
        for (int i = 0; i < kTpt; ++i) {
            const int t = tid * kTpt + i;
            cnt[t] = run;
            wb[t] = cur[i] - run;  // modular: wb[t] + slot is the entry's sorted position
            cur[i] += v[i];
            run += v[i];
        }
        const uint32_t nsub = total;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPer; ++k) {  // counting sort into LDS
            if (rr[k] != 0xFFFFFFFFu) {
                const uint32_t slot = cnt[x.c[k] >> p.tile_shift] + rr[k];
                s_cell[slot] = x.c[k];
                if constexpr (HAS_VAL) s_val[slot] = x.v[k];
            }
        }
        __syncthreads();
        // a fresh copy of tid per round: without it the compiler hoists all kPer slot
        // indices out of the round loop and spills registers to hold them
        int32_t me;
        asm volatile("v_mov_b32 %0, %1" : "=v"(me) : "v"(tid));
        // write out: kPer stores per thread, always; kOut entries at a time, their LDS
        // reads issued together (one entry at a time left each store waiting on two
        // dependent LDS round trips)
        constexpr int kOut = 4;
        static_assert(kPer % kOut == 0, "write-out batches");
#pragma unroll
        for (int k0 = 0; k0 < kPer; k0 += kOut) {
            uint32_t c[kOut], w[kOut];
            TV vv[kOut];
#pragma unroll
            for (int b = 0; b < kOut; ++b) {
                const uint32_t j = (uint32_t)((k0 + b) * kPartThreads + me);
                c[b] = s_cell[j];
                if constexpr (HAS_VAL) vv[b] = s_val[j];
            }
#pragma unroll
            for (int b = 0; b < kOut; ++b) {
                const bool ok = me < (int32_t)nsub - (k0 + b) * kPartThreads;
                w[b] = wb[ok ? c[b] >> p.tile_shift : 0u];
            }
#pragma unroll
            for (int b = 0; b < kOut; ++b) {
                const uint32_t j = (uint32_t)((k0 + b) * kPartThreads + me);
                const bool ok = me < (int32_t)nsub - (k0 + b) * kPartThreads;
                const size_t dst = ok ? (size_t)(w[b] + j) : sink;
                p.sorted_cell[dst] = c[b];
                if constexpr (HAS_VAL) sval[dst] = vv[b];
            }
            __builtin_amdgcn_sched_barrier(0);  // bounded registers
        }
#pragma unroll
        for (int i = 0; i < kTpt; ++i) cnt[tid * kTpt + i] = 0u;
        __syncthreads();
    };
    R ra, rb;
    load(ra, 0);
    for (int32_t r = 0; r < (int32_t)rounds; r += 2) {  // an odd count ends with an empty round
        load(rb, r + 1);
        round(ra, r);
        load(ra, r + 2);
        round(rb, r + 1);
    }
}

// ---- per-tile aggregation of the sorted spill ------------------------------------------------
// Entries [b, e) of one tile's sorted spill, kU per thread per step, loaded two steps
// ahead with unconditional loads (an index past the end re-reads entry b and is dropped).
template <typename TV, bool HAS_VAL>
struct TileRows {
    static constexpr int kU = 4;
    uint32_t gc[kU];
    TV vv[kU];

    __device__ __forceinline__ void load(const PartArgs &p, unsigned long long j0,
                                         unsigned long long b, unsigned long long e) {
        const TV *val = reinterpret_cast<const TV *>(p.sorted_val);
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            unsigned long long j = j0 + (unsigned long long)(u * kTileThreads + threadIdx.x);
            j = j < e ? j : b;
            gc[u] = p.sorted_cell[j];
            if constexpr (HAS_VAL) vv[u] = val[j];
            else vv[u] = TV(0);
        }
    }
};

// Write tile t's cells from its LDS region: every cell outside the window (the window
// merge wrote the others; their only spilled rows are rows with a NaN value when count
// and count(v) are both asked for, and those are added).
template <int KEYB>
__device__ void tile_store(const uint8_t *smem, const Lay &L, const PartArgs &p, int t,
                           const int32_t *window, int32_t W, int32_t H, int64_t wcap_px,
                           int32_t ncat, int64_t ncells, const Planes &out) {
    const int64_t c0 = (int64_t)t * p.tile_cells;
    const Window w = load_window(window, W, H, wcap_px);
    int64_t c1 = c0 + p.tile_cells;
    if (c1 > ncells) c1 = ncells;
    for (int64_t gcell = c0 + threadIdx.x; gcell < c1; gcell += kTileThreads) {
        // cells < 2^32 (checked on the host): 32-bit divisions, not 64-bit ones
        const uint32_t gc = (uint32_t)gcell;
        const uint32_t pix = ncat == 1 ? gc : gc / (uint32_t)ncat;
        const int32_t py = (int32_t)(pix / (uint32_t)W), px = (int32_t)(pix - (uint32_t)py * (uint32_t)W);
        if (in_window(w, px, py)) {
            if (L.flags & PL_NAN) region_add_to<KEYB>(smem, L, (uint32_t)(gcell - c0), out, gcell);
            continue;
        }
        region_store<KEYB>(smem, L, (uint32_t)(gcell - c0), out, gcell);
    }
}

// WF: the plane set when known at compile time (0: read from the layout).
template <typename TV, bool HAS_VAL, int KEYB, bool SQDEV, uint32_t WF = 0>
__global__ __launch_bounds__(kTileThreads) void tile_agg_kernel(PartArgs p, Lay L,
                                                               const int32_t *window, int32_t W,
                                                               int32_t H, int64_t wcap_px,
                                                               int32_t ncat, int64_t ncells,
                                                               const double *mean, Planes out) {
    if (tiles_run(window)) return;  // uniform: the tile strategy's tile pass runs instead
    extern __shared__ __align__(16) uint8_t smem[];
    using R = TileRows<TV, HAS_VAL>;
    constexpr unsigned long long kStep = (unsigned long long)R::kU * kTileThreads;
    // this workgroup's part: the last tile whose first part is at or before it
    const uint32_t part = blockIdx.x;
    if (part >= p.part_start[p.ntiles] || part >= (uint32_t)p.max_parts) return;  // the grid is an upper bound
    const int t = (int)p.part_tile[part];
    region_init<KEYB>(smem, L, threadIdx.x, kTileThreads);
    __syncthreads();
    const uint32_t first = p.part_start[t], nparts = p.part_start[t + 1] - first;
    const unsigned long long tb = p.tile_start[t], te = p.tile_start[t + 1];
    const unsigned long long len = (te - tb + nparts - 1) / nparts;  // an even split
    unsigned long long b = tb + (unsigned long long)(part - first) * len;
    unsigned long long e = b + len;
    b = b < te ? b : te;
    e = e < te ? e : te;
    const int64_t c0 = (int64_t)t * p.tile_cells;
    if (b < e) {
        // three register sets rotate roles, as in the scan
        R ra, rb, rc;
        ra.load(p, b, b, e);
        rb.load(p, b + kStep, b, e);
        unsigned long long j0 = b;
        auto step = [&](R &cur, R &far) {
            far.load(p, j0 + 2 * kStep, b, e);
#pragma unroll
            for (int u = 0; u < R::kU; ++u) {
                if (j0 + (unsigned long long)(u * kTileThreads + threadIdx.x) >= e) continue;
                const uint32_t lc = (uint32_t)(cur.gc[u] - c0);
                if (lc >= L.cap) continue;  // never, for entries the scatter wrote
                double v = 0.0;
                bool vnan = true;
                if constexpr (HAS_VAL) {
                    v = (double)cur.vv[u];
                    vnan = v != v;
                }
                double add = v;
                if constexpr (SQDEV) {
                    const double d = v - mean[cur.gc[u]];
                    add = d * d;
                }
                region_add<KEYB, WF>(smem, L, lc, vnan, v, add);
            }
            j0 += kStep;
        };
        while (j0 < e) {  // steps past e load entry b and add nothing
            step(ra, rc);
            step(rb, ra);
            step(rc, rb);
        }
    }
    __syncthreads();
    if (nparts > 1) {
        // a split tile: every part leaves its planes in its scratch region and
        // tile_merge_kernel, launched next, folds them and writes the tile (a kernel
        // boundary publishes the regions: a device-scope fence per part, as a last-part
        // merge needs, writes back L2 each time and cost more than the tile work)
        uint4 *mine = reinterpret_cast<uint4 *>(p.scratch + (size_t)part * L.bytes);
        const uint4 *src = reinterpret_cast<const uint4 *>(smem);
        for (uint32_t j = threadIdx.x; j < L.bytes / 16u; j += kTileThreads) mine[j] = src[j];
        return;
    }
    tile_store<KEYB>(smem, L, p, t, window, W, H, wcap_px, ncat, ncells, out);
}

// One plane of a split tile: cell c of the LDS region = op over the parts' scratch
// regions (part q at parts + q * part_bytes).  Each of the NT threads folds kFold cells
// at a time, so each part's loads for them are in flight together (an index past cap
// re-reads the last cell and is not stored).
template <int NT, typename T, typename Op>
__device__ __forceinline__ void fold_plane(uint8_t *lds, const uint8_t *parts, size_t part_bytes,
                                           uint32_t off, uint32_t nparts, uint32_t cap, T ident,
                                           Op op) {
    constexpr int kFold = 8;
    T *dst = reinterpret_cast<T *>(lds + off);
    for (uint32_t c0 = threadIdx.x; c0 < cap; c0 += kFold * NT) {
        T acc[kFold];
#pragma unroll
        for (int b = 0; b < kFold; ++b) acc[b] = ident;
        for (uint32_t q = 0; q < nparts; ++q) {
            const T *src = reinterpret_cast<const T *>(parts + (size_t)q * part_bytes + off);
            T v[kFold];
#pragma unroll
            for (int b = 0; b < kFold; ++b) {
                const uint32_t c = c0 + (uint32_t)(b * NT);
                v[b] = src[c < cap ? c : cap - 1];
            }
#pragma unroll
            for (int b = 0; b < kFold; ++b) acc[b] = op(acc[b], v[b]);
        }
#pragma unroll
        for (int b = 0; b < kFold; ++b) {
            const uint32_t c = c0 + (uint32_t)(b * NT);
            if (c < cap) dst[c] = acc[b];
        }
    }
}

// Every plane of a split tile's region folded into LDS.
template <int NT, int KEYB>
__device__ __forceinline__ void fold_parts(uint8_t *smem, const Lay &L, const uint8_t *parts,
                                           uint32_t nparts) {
    using T = typename KeyOf<KEYB>::T;
    auto add64 = [](double a, double b) { return a + b; };
    auto add32 = [](uint32_t a, uint32_t b) { return a + b; };
    auto kmin = [](T a, T b) { return b < a ? b : a; };
    auto kmax = [](T a, T b) { return b > a ? b : a; };
    if (L.flags & PL_SUM) fold_plane<NT, double>(smem, parts, L.bytes, L.o_sum, nparts, L.cap, 0.0, add64);
    if (L.flags & PL_MIN) fold_plane<NT, T>(smem, parts, L.bytes, L.o_min, nparts, L.cap, KeyOf<KEYB>::kMinEmpty, kmin);
    if (L.flags & PL_MAX) fold_plane<NT, T>(smem, parts, L.bytes, L.o_max, nparts, L.cap, T(0), kmax);
    if (L.flags & PL_CNT) fold_plane<NT, uint32_t>(smem, parts, L.bytes, L.o_cnt, nparts, L.cap, 0u, add32);
    if (L.flags & PL_CNTV) fold_plane<NT, uint32_t>(smem, parts, L.bytes, L.o_cntv, nparts, L.cap, 0u, add32);
    if (L.flags & PL_NAN) fold_plane<NT, uint32_t>(smem, parts, L.bytes, L.o_nan, nparts, L.cap, 0u, add32);
}

// Split tiles: one workgroup per tile folds its parts' scratch regions in LDS and writes
// the tile.  Tiles of one part were written by tile_agg_kernel.
template <int KEYB>
__global__ __launch_bounds__(kTileThreads) void tile_merge_kernel(PartArgs p, Lay L,
                                                                 const int32_t *window, int32_t W,
                                                                 int32_t H, int64_t wcap_px,
                                                                 int32_t ncat, int64_t ncells,
                                                                 Planes out) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int t = blockIdx.x;
    if (t >= p.ntiles) return;
    const uint32_t first = p.part_start[t], nparts = p.part_start[t + 1] - first;
    if (nparts <= 1) return;  // uniform
    fold_parts<kTileThreads, KEYB>(smem, L, p.scratch + (size_t)first * L.bytes, nparts);
    __syncthreads();
    tile_store<KEYB>(smem, L, p, t, window, W, H, wcap_px, ncat, ncells, out);
}

// ---- the tile strategy ---------------------------------------------------------------------
// With rows spread over the canvas (many categories, uniform data) a hot window holds few
// of them, and the window strategy spills nearly every row: 4 + value bytes written by
// the scan, read and written again by the scatter, read by the tile pass.  The tile
// strategy keeps no window.  Its scan gives every tile an LDS bucket of C = 2^cshift
// entries (a 16-bit tile-local cell offset, and the value) and writes a bucket that fills
// as one block of its workgroup's own block stream (nothing is reserved across
// workgroups); rows that arrive at a full bucket go straight into the next blocks of
// their tile.  At its end each scan workgroup lists its blocks in tile order, and the tile
// pass aggregates each tile's blocks in LDS as it does the spill list.  A row costs one
// write and one read of 2 + value bytes.
constexpr int kBktMinC = 16;              // smallest bucket worth a block
constexpr int kBktMaxC = 256;
constexpr uint32_t kBktMaxTileCells = 32768;  // tile offsets < 2^15: 0xFFFF marks unused slots
constexpr uint16_t kBktUnused = 0xFFFF;   // an unused slot of a partial block
constexpr int kBktBatch = 1024;           // block ids the tile pass stages in LDS at a time
constexpr int kBktBatchMax = 4096;        // (more, where two workgroups still fit a CU)
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

struct BktArgs {
    const int32_t *rec;            // the call's window record (its strategy)
    int32_t nb;                    // buckets (= tiles)
    uint32_t tile_shift, cshift;   // tile cells = 1 << tile_shift, block entries = 1 << cshift
    uint32_t o_bval;               // byte offset of a block's values
    uint32_t blk_bytes;
    uint32_t bcap;                 // blocks one scan workgroup may write
    uint8_t *pool;                 // [G * bcap] blocks
    unsigned long long *hdr;       // [G * bcap] tile | rank << 32 (among its workgroup's
                                   // blocks of that tile)
    uint32_t *nblk;                // [G] blocks each scan workgroup wrote
    uint32_t *hist;                // [G][nb] blocks of each tile
    uint32_t *wl;                  // [G][bcap]: each workgroup's block indices ordered by
                                   // tile (tile t's run starts at lpre[g][t])
    uint32_t *lpre;                // [G][nb]
    uint32_t o_cnt, o_lo, o_tblk, o_base, o_bk, o_fl, o_off, o_val, o_ctrl;  // scan LDS offsets
    uint32_t loc_cap;              // block indices the bucket area of LDS holds
    uint32_t batch;                // block ids the tile pass stages at a time
};

// The tile strategy's scan.  Bucket b's slots are numbered from its LDS block: slots
// [0, C) are the block's positions, and a flush that writes k blocks out takes k C off the
// count.  Each row takes the next slot of its tile's bucket (an LDS atomic); a row whose
// slot is past the LDS block keeps it in registers.  After every kSteps chunks (three
// for float32 coordinates without a value column, else one: the rows held cost registers)
// one barrier reads whether any row did, and if so the workgroup flushes: every bucket
// holding C slots or more writes its LDS block and the full blocks its overflowing rows
// fill (those rows write theirs directly), its count drops by the slots written, and the
// rows left over take their LDS positions (the next chunks' rows insert at later slots,
// whose positions cannot collide with theirs).  A flush is one more barrier and a few LDS round trips during which no
// row is binned, so they are batched: few checks, every LDS read of a phase issued
// together, the full blocks of a wave's buckets copied by all its lanes.  (A
// barrier-free form, in which the lane completing a block writes it out alone, measured
// slower: those copies run with one lane of a wave active.)  The buckets left at the end
// are written as partial blocks (unused slots marked).
constexpr int kBktOwn = 4;                      // buckets a thread owns (nb <= 4 * 1024)
static_assert(kBktOwn * kScanThreads >= kMaxTiles, "every bucket has an owner");
constexpr int kBktSpare = 64;                   // spare bucket words (one per lane)

template <typename TC, typename TV, bool HAS_VAL, int CK, bool LIN, bool WIDE = false>
__global__ __launch_bounds__(kScanThreads) void bscan_kernel(PointArgs a, BktArgs s) {
    if (!tiles_run(s.rec)) return;  // uniform: the window strategy's scan runs instead
    constexpr int kSteps = (HAS_VAL || sizeof(TC) == 8) ? 1 : 3;  // chunks between checks
    constexpr int kBktPend = kSteps * kItems;    // rows a thread may hold past their LDS block
    extern __shared__ __align__(16) uint8_t smem[];
    // per bucket (+ kBktSpare spare words, which dropped rows and idle lanes address)
    uint32_t *cnt = reinterpret_cast<uint32_t *>(smem + s.o_cnt);     // slots taken
    uint32_t *lo = reinterpret_cast<uint32_t *>(smem + s.o_lo);       // a flush: slots written
    uint32_t *bbase = reinterpret_cast<uint32_t *>(smem + s.o_base);  // a flush: block of slot v
                                                                      // is bbase + v / C
    uint32_t *xs = reinterpret_cast<uint32_t *>(smem + s.o_bk);       // end: a scan's scratch
                                                                      // (before: spare slots)
    uint32_t *tblk = reinterpret_cast<uint32_t *>(smem + s.o_tblk);   // blocks written
    uint2 *fl = reinterpret_cast<uint2 *>(smem + s.o_fl);             // end: (bucket, block)
    uint16_t *boff = reinterpret_cast<uint16_t *>(smem + s.o_off);    // [nb][C] offsets
    TV *bval = reinterpret_cast<TV *>(smem + s.o_val);                // [nb][C] values
    // [0] blocks written, [1] buckets listed, [2 + k % 3] the flag of check k
    uint32_t *ctrl = reinterpret_cast<uint32_t *>(smem + s.o_ctrl);
    const int tid = threadIdx.x, lane = tid & 63;
    const uint32_t cs = s.cshift, C = 1u << cs, cmask = C - 1u;
    const uint32_t omask = (1u << s.tile_shift) - 1u;
    const uint32_t spare = (uint32_t)s.nb + (uint32_t)lane;
    const int ncat = a.cat.ncat;
    const uint32_t wn = (uint32_t)a.W * (uint32_t)(CK == CK_NONE ? 1 : ncat);  // < 2^24 (host)
    // a one-byte code b (zero-extended) is category b - (b >= c8_at ? c8_sub : 0), dropped
    // unless that is below ncat: for uint8 codes c8_at = 256 (never); for int8 codes the
    // bytes >= 128 are k = b - 256 < 0, which the reference wraps once to k + ncat
    const uint32_t c8_at = a.cat.dtype == DSX_I8 ? 128u : 256u;
    const uint32_t c8_sub = 256u - (uint32_t)ncat;
    for (int b = tid; b < s.nb + kBktSpare; b += kScanThreads) {
        cnt[b] = 0u; lo[b] = 0u; bbase[b] = 0u;
        if (b < s.nb) tblk[b] = 0u;
    }
    if (tid < 5) ctrl[tid] = 0u;
    __syncthreads();
    const size_t g0 = (size_t)blockIdx.x * s.bcap;  // this workgroup's first block
    auto blk = [&](uint32_t q) { return s.pool + (g0 + q) * (size_t)s.blk_bytes; };
    // rows held past their bucket's LDS block: slot, tile << 16 | tile offset, value
    uint32_t pend[kBktPend], ptag[kBktPend];
    TV pv[kBktPend];
#pragma unroll
    for (int i = 0; i < kBktPend; ++i) { pend[i] = kNoSlot; ptag[i] = 0u; pv[i] = TV(0); }
    auto lds_at = [&](uint32_t t, uint32_t slot) { return (t << cs) | (slot & cmask); };
    // a flush's list of the wave's full buckets (bucket, first block) in the end-of-scan
    // list area, 64 entries a wave; log2 of the 16-byte pieces of a block's offsets and
    // of its values
    uint2 *const wfl = fl + (tid >> 6) * 64;
    // each lane's spare offset and value positions, in the end-of-scan scratch words
    const uint32_t off_spare = (s.o_bk - s.o_off) / 2u + (uint32_t)lane;
    const uint32_t val_spare = (s.o_bk - s.o_val) / (uint32_t)sizeof(TV) + (uint32_t)lane;
    const uint32_t lco = cs - 3u;
    const uint32_t lcv = HAS_VAL ? cs + (sizeof(TV) == 8 ? 3u : 2u) - 4u : 0u;
    auto flush = [&]() {  // every thread calls it, after a check's barrier (counts stable)
        // A. every bucket holding C slots or more gets k = count / C consecutive blocks of
        //    the stream and their headers; its wave copies its LDS block out
#pragma unroll
        for (int o = 0; o < kBktOwn; ++o) {
            if (o * kScanThreads >= s.nb) break;  // uniform
            const int b = tid + o * kScanThreads;
            uint32_t n = 0u, k = 0u;
            if (b < s.nb) {
                n = cnt[b];
                k = n >= C ? n >> cs : 0u;
            }
            // blocks: the lanes with one block take consecutive ones after one atomic for
            // the wave (a lane-mask prefix, no LDS round trip); a lane with more takes its own
            const unsigned long long m1 = __ballot(k == 1u);
            uint32_t q0 = 0u;
            if (m1) {  // uniform
                const int first = __ffsll((long long)m1) - 1;
                uint32_t b1 = 0u;
                if (lane == first) b1 = atomicAdd(&ctrl[0], (uint32_t)__popcll(m1));
                b1 = (uint32_t)__builtin_amdgcn_readlane((int)b1, first);
                q0 = b1 + __builtin_amdgcn_mbcnt_hi((uint32_t)(m1 >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)m1, 0u));
            }
            if (k >= 2u) q0 = atomicAdd(&ctrl[0], k);
            const unsigned long long mf = __ballot(k != 0u);
            if (k != 0u) {
                const uint32_t r0 = tblk[b];
                bbase[b] = q0;          // the block of a held row's slot v is bbase + v / C
                tblk[b] = r0 + k;
                lo[b] = k << cs;        // held rows below this slot go to those blocks
                cnt[b] = n - (k << cs); // (no row inserts during this phase)
                for (uint32_t q = 0; q < k; ++q)
                    if (q0 + q < s.bcap)  // always: bcap bounds a workgroup's blocks
                        s.hdr[g0 + q0 + q] = (unsigned long long)b | ((unsigned long long)(r0 + q) << 32);
                const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(mf >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)mf, 0u));
                wfl[r] = make_uint2((uint32_t)b, q0);
            }
            if (mf == 0ull) continue;  // uniform
            // the wave's full LDS blocks, copied by all its lanes 16 bytes each, consecutive
            // lanes along a block, so that every store writes whole lines
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const uint32_t nf = (uint32_t)__popcll(mf);
            for (uint32_t d = (uint32_t)lane; d < (nf << lco); d += 64u) {
                const uint2 e = wfl[d >> lco];
                if (e.y >= s.bcap) continue;
                const uint32_t w = d & ((1u << lco) - 1u);
                reinterpret_cast<uint4 *>(blk(e.y))[w] =
                    reinterpret_cast<const uint4 *>(boff + (e.x << cs))[w];
            }
            if constexpr (HAS_VAL) {
                for (uint32_t d = (uint32_t)lane; d < (nf << lcv); d += 64u) {
                    const uint2 e = wfl[d >> lcv];
                    if (e.y >= s.bcap) continue;
                    const uint32_t w = d & ((1u << lcv) - 1u);
                    reinterpret_cast<uint4 *>(blk(e.y) + s.o_bval)[w] =
                        reinterpret_cast<const uint4 *>(bval + (e.x << cs))[w];
                }
            }
            // the list is rewritten by the next bucket set's owners only after every lane's
            // reads above (one wave, in order)
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
        __syncthreads();
        // B. a row held goes to the full block of its slot (straight to the stream) when
        //    the slot is below k C, else to the front of the bucket's next LDS block
        //    (position slot % C)
        if constexpr (!HAS_VAL) {
            // a wave gathers its held rows (about one in ten) into its 64-entry list and
            // places them 64 at a time with every lane busy, instead of testing each of its
            // kBktPend registers in every lane
            uint32_t n = 0u;  // uniform: rows in the wave's list
            auto place = [&](uint32_t cnt) {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                if ((uint32_t)lane < cnt) {
                    const uint2 e = wfl[lane];  // (slot, tile << 16 | tile offset)
                    const uint32_t t = e.y >> 16;
                    const uint32_t bw = bbase[t], lw = lo[t];
                    // (two store sites: one that could write either memory would be a flat
                    // store, whose LDS waits would also wait for its global write)
                    if (e.x < lw) {
                        const uint32_t qb = bw + (e.x >> cs);
                        if (qb < s.bcap)
                            global_ptr(reinterpret_cast<uint16_t *>(blk(qb)))[e.x & cmask] = (uint16_t)e.y;
                    } else {
                        boff[lds_at(t, e.x)] = (uint16_t)e.y;
                    }
                }
                // the list is rewritten only after every lane's read above (one wave, in order)
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            };
#pragma unroll
            for (int i = 0; i < kBktPend; ++i) {
                const bool h = pend[i] != kNoSlot;
                const unsigned long long m = __ballot(h);
                if (m == 0ull) continue;  // uniform
                const uint32_t c = (uint32_t)__popcll(m);
                if (n + c > 64u) { place(n); n = 0u; }
                if (h) {
                    const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                    wfl[n + r] = make_uint2(pend[i], ptag[i]);
                }
                n += c;
                pend[i] = kNoSlot;
            }
            if (n != 0u) place(n);
        } else {
            uint32_t bw[kBktPend], lw[kBktPend];
#pragma unroll
            for (int i = 0; i < kBktPend; ++i) {
                const uint32_t idx = pend[i] != kNoSlot ? ptag[i] >> 16 : spare;
                bw[i] = bbase[idx];
                lw[i] = lo[idx];
            }
            // (two loops: one store site that could write either memory would be a flat
            // store, and the LDS waits after it would also wait for its write to global memory)
#pragma unroll
            for (int i = 0; i < kBktPend; ++i) {
                if (pend[i] == kNoSlot || pend[i] >= lw[i]) continue;
                const uint32_t qb = bw[i] + (pend[i] >> cs);
                if (qb < s.bcap) {
                    uint8_t *dst = blk(qb);
                    reinterpret_cast<uint16_t *>(dst)[pend[i] & cmask] = (uint16_t)ptag[i];
                    reinterpret_cast<TV *>(dst + s.o_bval)[pend[i] & cmask] = pv[i];
                }
                pend[i] = kNoSlot;
            }
#pragma unroll
            for (int i = 0; i < kBktPend; ++i) {
                if (pend[i] == kNoSlot) continue;
                const uint32_t at = lds_at(ptag[i] >> 16, pend[i]);
                boff[at] = (uint16_t)ptag[i];
                bval[at] = pv[i];
                pend[i] = kNoSlot;
            }
        }
        // no barrier: the next chunks' rows take slots past every one just placed (phase A
        // lowered the counts before the barrier), so no LDS position collides; the next
        // check's barrier orders these writes before the next flush
    };
    const int64_t nchunks = (a.n + kChunk - 1) / kChunk;
    using R = Rows<TC, TV, HAS_VAL, CK, WIDE>;
    // chunks loaded ahead: three where the checks come every three chunks (four register
    // sets: the loads in flight cover a flush), else two; the generic category column (a
    // 64-bit raw value a row) keeps two, as a fourth set would spill registers
    constexpr int kAhead = (kSteps == 3 && CK != CK_ANY) ? 3 : 2;
    R ra, rb, rc, rd;
    ra.load(a, blockIdx.x, tid);
    rb.load(a, (int64_t)blockIdx.x + gridDim.x, tid);
    if constexpr (kAhead == 3) rc.load(a, (int64_t)blockIdx.x + 2 * (int64_t)gridDim.x, tid);
    int64_t ck = blockIdx.x;
    uint32_t fk = 0;  // this check's flag: ctrl[2 + fk]
    bool over = false;
    // pb: the first of this chunk's kItems entries of pend (a constant at every call)
    auto step = [&](R &cur, R &far, const int pb) {
        far.load(a, ck + kAhead * (int64_t)gridDim.x, tid);
        const int64_t rem64 = a.n - ck * kChunk;
        const uint32_t rem = rem64 >= kChunk ? (uint32_t)kChunk
                                             : (rem64 > 0 ? (uint32_t)rem64 : 0u);
        // three passes over the chunk's rows, so that the LDS round trips of all its rows
        // are in flight together: bin every row (a dropped row takes its lane's spare
        // bucket, whose slots are never stored), take every slot, then store
        uint32_t tt[kItems], ss[kItems];
        uint16_t oo[kItems];
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            const uint32_t j = WIDE ? (uint32_t)(4 * tid + it) : (uint32_t)(it * kScanThreads + tid);
            int32_t px = 0, py = 0, pc = 0;
            bool keep;
            if constexpr (LIN) {  // as in scan_kernel, without its selects (cvt_sat_i32)
                const double x = (double)cur.x[it], y = (double)cur.y[it];
                const bool inr = (a.xmin <= x) & (x <= a.xmax) & (a.ymin <= y) & (y <= a.ymax);
                const double fx = __dadd_rn(__dmul_rn(x, a.sx), a.tx);
                const double fy = __dadd_rn(__dmul_rn(y, a.sy), a.ty);
                const int32_t xx = cvt_sat_i32(fx), yy = cvt_sat_i32(fy);
                px = xx < a.W - 1 ? xx : a.W - 1;
                py = yy < a.H - 1 ? yy : a.H - 1;
                keep = (j < rem) & inr;
            } else {
                keep = j < rem && locate_xy<TC>(a, cur.x[it], cur.y[it], px, py);
            }
            if constexpr (CK == CK_CODE8) {
                const uint32_t b8 = (uint32_t)cur.c[it];
                const uint32_t k = b8 - (b8 >= c8_at ? c8_sub : 0u);
                keep = keep & (k < (uint32_t)ncat);
                pc = (int32_t)k;
            } else if constexpr (CK != CK_NONE) {
                pc = scan_category<CK>(a.cat, cur.c[it]);
                keep = keep && pc >= 0;
            }
            // cells < 2^32 (checked on the host), so 32-bit arithmetic is exact; py, W ncat,
            // px and ncat are < 2^24, so the cell is 24-bit multiply-adds
            const uint32_t cell = CK == CK_NONE
                ? __umul24((uint32_t)py, wn) + (uint32_t)px
                : __umul24((uint32_t)py, wn) + __umul24((uint32_t)px, (uint32_t)ncat) + (uint32_t)pc;
            tt[it] = keep ? cell >> s.tile_shift : spare;
            oo[it] = (uint16_t)(cell & omask);
        }
#pragma unroll
        for (int it = 0; it < kItems; ++it) ss[it] = atomicAdd(&cnt[tt[it]], 1u);
        // without branches: a row that does not go to its LDS block (dropped, or held:
        // its slot is past the block) writes its lane's spare position instead
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            const uint32_t t = tt[it];
            const bool keep = t < (uint32_t)s.nb;
            const bool inl = keep & (ss[it] < C);
            const bool held = keep & !inl;
            boff[inl ? lds_at(t, ss[it]) : off_spare] = oo[it];
            if constexpr (HAS_VAL) bval[inl ? lds_at(t, ss[it]) : val_spare] = cur.v[it];
            pend[pb + it] = held ? ss[it] : kNoSlot;
            ptag[pb + it] = (t << 16) | oo[it];
            if constexpr (HAS_VAL) pv[pb + it] = cur.v[it];
            over |= held;
        }
        ck += gridDim.x;
    };
    auto check = [&]() {
        if (over) ctrl[2 + fk] = 1u;
        __syncthreads();
        const bool any = ctrl[2 + fk] != 0u;
        // the previous check's flag: every wave read it before this barrier, and the
        // chunks that raise it next start after the next check's barrier
        if (tid == 0) ctrl[2 + (fk == 0u ? 2u : fk - 1u)] = 0u;
        if (any) flush();
        over = false;
        fk = fk == 2u ? 0u : fk + 1u;
    };
    if constexpr (kAhead == 3) {
        while (ck < nchunks) {  // the four sets rotate through twelve steps, four checks
            step(ra, rd, 0); step(rb, ra, kItems); step(rc, rb, 2 * kItems); check();
            step(rd, rc, 0); step(ra, rd, kItems); step(rb, ra, 2 * kItems); check();
            step(rc, rb, 0); step(rd, rc, kItems); step(ra, rd, 2 * kItems); check();
            step(rb, ra, 0); step(rc, rb, kItems); step(rd, rc, 2 * kItems); check();
        }
    } else {
        while (ck < nchunks) {  // three steps per trip, as in scan_kernel
            step(ra, rc, 0);
            if constexpr (kSteps == 1) check();
            step(rb, ra, kSteps == 1 ? 0 : kItems);
            if constexpr (kSteps == 1) check();
            step(rc, rb, kSteps == 1 ? 0 : 2 * kItems);
            check();
        }
    }
    __syncthreads();
    // the buckets left: one partial block each (every bucket holds < C slots)
    for (int b = tid; b < s.nb; b += kScanThreads) {
        const uint32_t n = cnt[b];
        if (n > 0u) {
            const uint32_t q0 = atomicAdd(&ctrl[0], 1u);
            if (q0 < s.bcap) s.hdr[g0 + q0] = (unsigned long long)b | ((unsigned long long)tblk[b] << 32);
            tblk[b] += 1u;
            fl[atomicAdd(&ctrl[1], 1u)] = make_uint2((uint32_t)b, q0);
        }
    }
    __syncthreads();
    const uint32_t nfl = ctrl[1];
    for (uint32_t d = tid; d < (nfl << cs); d += kScanThreads) {  // one entry per thread
        const uint2 e = fl[d >> cs];
        const uint32_t w = d & cmask, n = cnt[e.x];
        if (e.y >= s.bcap) continue;
        uint8_t *dst = blk(e.y);
        reinterpret_cast<uint16_t *>(dst)[w] = w < n ? boff[(e.x << cs) + w] : kBktUnused;
        if constexpr (HAS_VAL)
            if (w < n) reinterpret_cast<TV *>(dst + s.o_bval)[w] = bval[(e.x << cs) + w];
    }
    // this workgroup's blocks in tile order, for the tile pass: tile t's run starts at the
    // exclusive prefix of the block counts over the tiles before it, and a block's rank
    // places it within the run (lpre reuses the flush bookkeeping words, free now)
    uint32_t *lp = bbase;
    {
        constexpr int kPer = kBktOwn;  // tiles per thread: 4 * 1024 >= nb
        uint32_t v[kPer], sum = 0u;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int b = tid * kPer + i;
            v[i] = b < s.nb ? tblk[b] : 0u;
            sum += v[i];
        }
        uint32_t total;
        uint32_t run = block_exscan32(sum, xs, total);
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int b = tid * kPer + i;
            if (b < s.nb) {
                lp[b] = run;
                s.lpre[(size_t)blockIdx.x * s.nb + b] = run;
            }
            run += v[i];
        }
    }
    __syncthreads();
    const uint32_t nq = ctrl[0] < s.bcap ? ctrl[0] : s.bcap;
    // the list is built in the bucket area of LDS (free now) when it fits, then written out
    // in order: written straight to its scattered positions it moved ~9x its bytes
    uint32_t *loc = reinterpret_cast<uint32_t *>(smem);
    const bool in_lds = nq <= s.loc_cap;
    for (uint32_t q = tid; q < nq; q += kScanThreads) {
        const unsigned long long h = s.hdr[g0 + q];
        const uint32_t t = (uint32_t)h;
        if (t >= (uint32_t)s.nb) continue;  // never, for blocks this workgroup wrote
        const uint32_t at = lp[t] + (uint32_t)(h >> 32);
        if (in_lds) loc[at] = q;
        else s.wl[g0 + at] = q;
    }
    if (in_lds) {  // uniform
        __syncthreads();
        for (uint32_t i = tid; i < nq; i += kScanThreads) s.wl[g0 + i] = loc[i];
    }
    for (int b = tid; b < s.nb; b += kScanThreads) s.hist[(size_t)blockIdx.x * s.nb + b] = tblk[b];
    if (tid == 0) s.nblk[blockIdx.x] = nq;
}

// Groups of four consecutive entries [g0, g0 + kStep) of a batch of staged blocks, kU
// groups per thread; a group index past the batch re-reads group 0 and is dropped.  A
// lane reads a group's four 16-bit offsets in one 8-byte load (a wave: four blocks), so
// four times the bytes are in flight per load instruction.
template <typename TV, bool HAS_VAL>
struct BktRows {
    static constexpr int kU = 4;
    uint2 o[kU];
    TV v[kU][4];

    __device__ __forceinline__ void load(const BktArgs &s, const uint32_t *sblk, uint32_t g0,
                                         uint32_t ngrp) {
        const uint32_t gs = s.cshift - 2u, gmask = (1u << gs) - 1u;  // groups in a block
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            uint32_t g = g0 + (uint32_t)(u * kTileThreads + threadIdx.x);
            g = g < ngrp ? g : 0u;
            const uint8_t *b = s.pool + (size_t)sblk[g >> gs] * s.blk_bytes;
            const uint32_t q = g & gmask;
            o[u] = reinterpret_cast<const uint2 *>(b)[q];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if constexpr (HAS_VAL) v[u][e] = reinterpret_cast<const TV *>(b + s.o_bval)[4u * q + e];
                else v[u][e] = TV(0);
            }
        }
    }
};

// The tile strategy's tile pass: part `blockIdx.x` of tile t aggregates its share of
// the tile's blocks in LDS (a split tile leaves its planes to tile_merge_kernel).
template <typename TV, bool HAS_VAL, int KEYB, bool SQDEV>
__global__ __launch_bounds__(kTileThreads) void btile_agg_kernel(PartArgs p, BktArgs s, Lay L,
                                                                const int32_t *window, int32_t W,
                                                                int32_t H, int64_t wcap_px,
                                                                int32_t ncat, int64_t ncells,
                                                                const double *mean, Planes out) {
    if (!tiles_run(window)) return;  // uniform: the window strategy's tile pass runs instead
    extern __shared__ __align__(16) uint8_t smem[];
    uint32_t *sblk = reinterpret_cast<uint32_t *>(smem + L.bytes);  // [s.batch]
    // the tile's blocks, in the order of their scan workgroups: workgroup g's run is
    // [ex[g], ex[g + 1]) of the tile's blocks (ex from spill_colscan_kernel), found in its
    // tile-ordered list at lp[g]
    uint32_t *ex = sblk + s.batch;                // [G + 1]
    uint32_t *lpg = ex + p.ngroups + 1;           // [G]
    using R = BktRows<TV, HAS_VAL>;
    constexpr uint32_t kStep = (uint32_t)R::kU * kTileThreads;
    const uint32_t part = blockIdx.x;
    if (part >= p.part_start[p.ntiles] || part >= (uint32_t)p.max_parts) return;  // the grid is an upper bound
    const int t = (int)p.part_tile[part];
    region_init<KEYB>(smem, L, threadIdx.x, kTileThreads);
    const uint32_t first = p.part_start[t], nparts = p.part_start[t + 1] - first;
    const unsigned long long tb = p.tile_start[t], te = p.tile_start[t + 1];
    const unsigned long long len = (te - tb + nparts - 1) / nparts;  // an even split of blocks
    unsigned long long b = tb + (unsigned long long)(part - first) * len;
    unsigned long long e = b + len;
    b = b < te ? b : te;
    e = e < te ? e : te;
    const int64_t c0 = (int64_t)t * p.tile_cells;
    const int G = p.ngroups;
    for (int g = threadIdx.x; g <= G; g += kTileThreads) {
        ex[g] = g < G ? p.hist[(size_t)g * p.ntiles + t] : (uint32_t)(te - tb);
        if (g < G) lpg[g] = s.lpre[(size_t)g * s.nb + t];
    }
    for (unsigned long long bb = b; bb < e; bb += s.batch) {
        const uint32_t nbb = (uint32_t)(e - bb < (unsigned long long)s.batch ? e - bb : s.batch);
        __syncthreads();  // the region and ex / lpg are written, the last batch's ids are read
        // each thread stages a run of consecutive ids: one binary search for the first
        // (the last workgroup whose run starts at or before it), then a forward walk
        const uint32_t per = (nbb + kTileThreads - 1) / kTileThreads;
        const uint32_t i0 = threadIdx.x * per;
        if (i0 < nbb) {
            uint32_t v = (uint32_t)(bb - tb) + i0;  // the block's index in the tile
            int g = 0, hi_ = G;
            while (hi_ - g > 1) {
                const int mid = (g + hi_) >> 1;
                if (ex[mid] <= v) g = mid;
                else hi_ = mid;
            }
            const uint32_t i1 = i0 + per < nbb ? i0 + per : nbb;
            for (uint32_t i = i0; i < i1; ++i, ++v) {
                while (g + 1 < G && ex[g + 1] <= v) ++g;  // (empty runs are skipped)
                const size_t gb = (size_t)g * s.bcap;
                sblk[i] = (uint32_t)(gb + s.wl[gb + lpg[g] + (v - ex[g])]);
            }
        }
        __syncthreads();
        const uint32_t ngrp = nbb << (s.cshift - 2u);  // groups of four entries
        R ra, rb, rc;  // three register sets rotate roles, as in the scan
        ra.load(s, sblk, 0u, ngrp);
        rb.load(s, sblk, kStep, ngrp);
        uint32_t g0 = 0u;
        auto step = [&](R &cur, R &far) {
            far.load(s, sblk, g0 + 2u * kStep, ngrp);
#pragma unroll
            for (int u = 0; u < R::kU; ++u) {
                if (g0 + (uint32_t)(u * kTileThreads + threadIdx.x) >= ngrp) continue;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const uint32_t lc = ((e < 2 ? cur.o[u].x : cur.o[u].y) >> (16 * (e & 1))) & 0xFFFFu;
                    if (lc >= L.cap) continue;  // an unused slot of a partial block
                    double v = 0.0;
                    bool vnan = true;
                    if constexpr (HAS_VAL) {
                        v = (double)cur.v[u][e];
                        vnan = v != v;
                    }
                    double add = v;
                    if constexpr (SQDEV) {
                        const double d = v - mean[c0 + lc];
                        add = d * d;
                    }
                    region_add<KEYB>(smem, L, lc, vnan, v, add);
                }
            }
            g0 += kStep;
        };
        while (g0 < ngrp) {  // steps past ngrp load group 0 and add nothing
            step(ra, rc);
            step(rb, ra);
            step(rc, rb);
        }
    }
    __syncthreads();
    if (nparts > 1) {  // a split tile, as in tile_agg_kernel
        uint4 *mine = reinterpret_cast<uint4 *>(p.scratch + (size_t)part * L.bytes);
        const uint4 *src = reinterpret_cast<const uint4 *>(smem);
        for (uint32_t j = threadIdx.x; j < L.bytes / 16u; j += kTileThreads) mine[j] = src[j];
        return;
    }
    tile_store<KEYB>(smem, L, p, t, window, W, H, wcap_px, ncat, ncells, out);
}

// ---- host side --------------------------------------------------------------------------------
struct Plan {
    uint32_t flags;
    int keyb, vbytes;
    Lay wl;              // window layout
    Lay tl;              // tile layout
    int64_t ncells;
    int32_t ntiles;
    int scan_groups;
    int64_t seg_cap;
    size_t scan_lds, tile_lds;
    uint32_t o_thist, o_stage_cell, o_stage_val;  // scan LDS offsets
    // workspace offsets
    size_t o_window, o_hist_sel, o_count, o_partials, o_spill_cell, o_spill_val, o_hist,
        o_totals, o_tile_start, o_sorted_cell, o_sorted_val, o_part_start, o_part_len, o_part_tile,
        o_scratch, total;
    int32_t max_parts;   // upper bound of the tile parts (grid of the tile kernel)
    bool whole;          // the window is the whole canvas
    bool global;         // the canvas is too large for the tiles: global-atomic pass
    int bshift, cw, ch;
    int tshift;          // tile_cells = 1 << tshift
    // the tile strategy (tiles_ok: the tiles' buckets fit the scan's LDS)
    bool tiles_ok;
    bool run_window, run_tiles;  // the strategies this call launches (both under AUTO)
    uint32_t cshift, o_bval, blk_bytes, bcap, min_part_tiles;
    uint32_t bo_cnt, bo_lo, bo_tblk, bo_base, bo_bk, bo_fl, bo_off, bo_val, bo_ctrl;  // its scan's LDS
    size_t bscan_lds, btile_lds;
    uint32_t batch;  // the tile pass's staged block ids
    size_t o_nblk, o_bhdr, o_bwl, o_blpre, o_pool;
};

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

static int make_plan(const dsx_points_desc *d, const dsx_view *v, const dsx_planes *pl, Plan &P) {
    if (plane_flags(pl) == 0) return DSX_ERR_ACCUM;
    P.flags = internal_flags(plane_flags(pl));
    const bool has_val = has_value(d);
    P.vbytes = has_val ? (d->val_dtype == DSX_F32 ? 4 : 8) : 0;
    P.keyb = (has_val && d->val_dtype == DSX_F32) ? 4 : 8;
    const int ncat = d->cat.kind == DSX_CAT_NONE ? 1 : d->cat.ncat;
    P.ncells = (int64_t)v->width * v->height * ncat;
    if (P.ncells >= (1ll << 32)) return DSX_ERR_SHAPE;
    const uint32_t cb = cell_bytes(P.flags, P.keyb);
    // tiles: runs of 2^tshift cells (a cell's tile is one shift) whose planes fit kTileLds
    // (several tile workgroups share a CU), or up to the whole LDS when that would make
    // more than kMaxTiles tiles
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && c > 0)
            cus = c;
    }
    P.tshift = 0;
    while ((int64_t)(2ll << P.tshift) * cb <= kTileLds - 16) ++P.tshift;
    // a small canvas still gets one tile per CU (down to 1024-cell tiles): 800x600 counts
    // made 30 tiles of 16384 cells, so 30 workgroups did all of the tile work
    while (P.tshift > kMinTileShift && ((P.ncells + (1ll << P.tshift) - 1) >> P.tshift) < cus)
        --P.tshift;
    while ((P.ncells + (1ll << P.tshift) - 1) >> P.tshift > kMaxTiles) ++P.tshift;
    const int64_t tcap = 1ll << P.tshift;
    P.global = false;
    if (tcap * (int64_t)cb > kLdsBytes - 16) {
        // more cells than kMaxTiles tiles of LDS hold (e.g. 8192 x 8192 count + mean):
        // one pass of global atomics instead, slower but exact
        P.global = true;
        // a finalized mean accumulates its sum and count(v) in the workspace unless the
        // caller asked for those planes too
        P.total = 256;
        if (pl->mean && !pl->sum) P.total += align_up(8 * (size_t)P.ncells, 256);
        if (pl->mean && !pl->countv) P.total += align_up(4 * (size_t)P.ncells, 256);
        return DSX_OK;
    }
    if (tcap % kTileThreads) return DSX_ERR_SHAPE;  // never: tiles hold >= 1024 cells
    P.ntiles = (int32_t)((P.ncells + tcap - 1) >> P.tshift);
    if (P.ntiles < 1) P.ntiles = 1;
    P.tl = make_lay(P.flags, P.keyb, (uint32_t)tcap);
    P.tile_lds = P.tl.bytes;
    // scan LDS: window planes, tile histogram, per-wave spill staging, control words.
    // The window keeps no NaN plane: with count() and count(v) both asked for, rows with
    // a NaN value are spilled and the tile adds them to the window's cells.
    const uint32_t wflags = P.flags & ~PL_NAN;
    const uint32_t cbw = cell_bytes(wflags, P.keyb);
    const size_t stage = (size_t)kWaves * kStage * (4 + (size_t)P.vbytes);
    const size_t side = align_up((size_t)4 * P.ntiles, 16) + stage + kCtrlBytes;
    const int64_t wcap_cells = (int64_t)((kLdsBytes - side - 16) / cbw);
    int64_t wcap_px = wcap_cells / ncat;
    if (wcap_px < 1) return DSX_ERR_SHAPE;
    P.whole = (int64_t)v->width * v->height <= wcap_px;
    if (P.whole) wcap_px = (int64_t)v->width * v->height;
    P.wl = make_lay(wflags, P.keyb, (uint32_t)(wcap_px * ncat));
    P.o_thist = P.wl.bytes;
    P.o_stage_cell = P.o_thist + (uint32_t)align_up((size_t)4 * P.ntiles, 16);
    P.o_stage_val = P.o_stage_cell + (uint32_t)(kWaves * kStage * 4);
    P.scan_lds = P.wl.bytes + side;
    // coarse selection grid: bins of 2^bshift pixels on a side, at most kSelBins bins
    P.bshift = 0;
    while ((int64_t)((v->width + (1 << P.bshift) - 1) >> P.bshift) *
               ((v->height + (1 << P.bshift) - 1) >> P.bshift) > kSelBins)
        ++P.bshift;
    P.cw = (v->width + (1 << P.bshift) - 1) >> P.bshift;
    P.ch = (v->height + (1 << P.bshift) - 1) >> P.bshift;
    // scan workgroups: one per CU at most, fewer for small inputs
    const int64_t chunks = (d->n + kChunk - 1) / kChunk;
    P.scan_groups = (int)(chunks < cus ? (chunks < 1 ? 1 : chunks) : cus);
    const int64_t group_rows = ((chunks + P.scan_groups - 1) / P.scan_groups) * (int64_t)kChunk;
    // each wave's spill sub-segment holds every row that wave could spill
    P.seg_cap = group_rows / kWaves;
    if (P.seg_cap < 1) P.seg_cap = 1;
    const size_t nseg = (size_t)P.scan_groups * kWaves * (size_t)P.seg_cap;
    // the tile strategy: the largest bucket C (a power of two, kBktMinC..kBktMaxC) whose
    // nb buckets fit the scan's LDS next to seven words per bucket
    P.tiles_ok = false;
    P.batch = kBktBatch;
    // (its scan forms a cell with 24-bit multiplies: py * (W ncat) + px * ncat + category)
    if (tcap <= (int64_t)kBktMaxTileCells && (int64_t)v->width * ncat < (1ll << 24)) {
        const size_t vb = (size_t)P.vbytes, nb = (size_t)P.ntiles;
        for (int cs = 8; (1 << cs) >= kBktMinC && !P.tiles_ok; --cs) {
            if ((1 << cs) > kBktMaxC) continue;
            const size_t C = (size_t)1 << cs;
            size_t q = 0;
            P.bo_val = (uint32_t)q; q = align_up(q + vb * C * nb, 16);   // values first (8-aligned)
            P.bo_off = (uint32_t)q; q = align_up(q + 2 * C * nb, 16);
            P.bo_cnt = (uint32_t)q; q = align_up(q + 4 * (nb + kBktSpare), 16);
            P.bo_lo = (uint32_t)q; q = align_up(q + 4 * (nb + kBktSpare), 16);
            P.bo_tblk = (uint32_t)q; q = align_up(q + 4 * nb, 16);
            P.bo_base = (uint32_t)q; q = align_up(q + 4 * (nb + kBktSpare), 16);
            // (the end's scan scratch, >= 17 words; during the scan, a spare offset and
            // value position for each lane of a wave)
            P.bo_bk = (uint32_t)q; q = align_up(q + 8 * 64, 16);
            // (the end-of-scan list; during the scan, the flushes' per-wave lists)
            P.bo_fl = (uint32_t)q; q = align_up(q + 8 * (nb > (size_t)kScanThreads ? nb : (size_t)kScanThreads), 16);
            P.bo_ctrl = (uint32_t)q; q += 32;
            const int64_t bcap = (group_rows + (int64_t)C - 1) / (int64_t)C + (int64_t)nb;
            // (512 bytes left for the kernel's static LDS words)
            if (q <= (size_t)kLdsBytes - 512 && (int64_t)P.scan_groups * bcap < (1ll << 32)) {
                P.tiles_ok = true;
                P.cshift = (uint32_t)cs;
                P.bscan_lds = q;
                P.bcap = (uint32_t)bcap;
                P.o_bval = (uint32_t)align_up(2 * C, 16);
                P.blk_bytes = (uint32_t)align_up(P.o_bval + vb * C, 16);
                const uint32_t mp = (uint32_t)(kMinPart >> cs);
                P.min_part_tiles = mp < 16u ? 16u : mp;
                // the tile pass stages kBktBatch block ids at a time, or up to
                // kBktBatchMax where two of its workgroups still fit a CU: each batch
                // starts its load pipeline over
                P.batch = kBktBatch;
                auto blds = [&](size_t nbat) {
                    return (size_t)P.tl.bytes + 4 * (nbat + 2 * (size_t)P.scan_groups + 1);
                };
                while (P.batch < (uint32_t)kBktBatchMax && blds(2 * (size_t)P.batch) <= (size_t)kLdsBytes / 2)
                    P.batch *= 2;
                P.btile_lds = blds(P.batch);
            }
        }
    }
    // which strategies run: the whole canvas in the window needs no choice; a forced
    // strategy runs alone; AUTO launches both and the kernels of the one the window pick
    // did not choose return at once
    const int strat = d->strategy;
    P.run_tiles = P.tiles_ok && !P.whole && strat != DSX_STRATEGY_WINDOW;
    P.run_window = !(P.run_tiles && strat == DSX_STRATEGY_TILES);
    // workspace: the shared part, then each strategy's buffers over the same bytes
    const int64_t n = d->n < 1 ? 1 : d->n;
    size_t o = 0;
    P.o_window = o; o = align_up(o + 4 * kRecWords, 256);
    P.o_hist_sel = o; o = align_up(o + 4 * (size_t)kSelBins, 256);
    P.o_count = o; o = align_up(o + 4 * (size_t)P.scan_groups * kWaves, 256);  // sub-segment lengths
    P.o_hist = o; o = align_up(o + 4 * (size_t)P.scan_groups * P.ntiles, 256);
    P.o_totals = o; o = align_up(o + 8 * (size_t)P.ntiles, 256);
    P.o_tile_start = o; o = align_up(o + 8 * (size_t)(P.ntiles + 1), 256);
    P.max_parts = 2 * P.ntiles + 1;
    P.o_part_start = o; o = align_up(o + 4 * (size_t)(P.ntiles + 1), 256);
    P.o_part_len = o; o = align_up(o + 4, 256);
    P.o_part_tile = o; o = align_up(o + 4 * (size_t)P.max_parts, 256);
    P.o_scratch = o; o = align_up(o + (size_t)P.max_parts * P.tl.bytes, 256);
    const size_t shared = o;
    P.o_partials = o; o = align_up(o + (size_t)P.scan_groups * P.wl.bytes, 256);
    P.o_spill_cell = o; o = align_up(o + 4 * (nseg + kScatterPad), 256);
    P.o_spill_val = o; o = align_up(o + (size_t)P.vbytes * (nseg + kScatterPad), 256);
    const size_t nsorted = (size_t)n + 64 * (size_t)P.scan_groups;  // + the scatter's sink lines
    P.o_sorted_cell = o; o = align_up(o + 4 * nsorted, 256);
    P.o_sorted_val = o; o = align_up(o + (size_t)P.vbytes * nsorted, 256);
    const size_t end_window = o;
    size_t end_tiles = shared;
    if (P.tiles_ok) {
        const size_t nblocks = (size_t)P.scan_groups * P.bcap;
        o = shared;
        P.o_nblk = o; o = align_up(o + 4 * (size_t)P.scan_groups, 256);
        P.o_bhdr = o; o = align_up(o + 8 * nblocks, 256);
        P.o_bwl = o; o = align_up(o + 4 * nblocks, 256);
        P.o_blpre = o; o = align_up(o + 4 * (size_t)P.scan_groups * P.ntiles, 256);
        P.o_pool = o; o = align_up(o + (size_t)P.blk_bytes * nblocks, 256);
        end_tiles = o;
    }
    P.total = shared;
    if (P.run_window && end_window > P.total) P.total = end_window;
    if (P.run_tiles && end_tiles > P.total) P.total = end_tiles;
    return DSX_OK;
}

template <typename TC, typename TV, bool HAS_VAL, int KEYB, bool SQDEV>
static int run_points(const dsx_points_desc *d, const PointArgs &a, const Plan &P,
                      const Planes &out, const double *mean, uint8_t *ws, hipStream_t st) {
    // Every kernel of this call reads the window from the call's own workspace: a stored
    // window (d->window) is copied in first, so that calls on other streams that share
    // the store and rewrite it cannot hand one call's kernels two different windows.
    int32_t *window = reinterpret_cast<int32_t *>(ws + P.o_window);
    uint32_t *hist_sel = reinterpret_cast<uint32_t *>(ws + P.o_hist_sel);
    uint32_t *seg_count = reinterpret_cast<uint32_t *>(ws + P.o_count);
    const int ncat = a.cat.ncat;
    const int64_t wcap_px = (int64_t)(P.wl.cap / (uint32_t)ncat);
    int s;
    if (P.whole) {
        hipLaunchKernelGGL(set_window_kernel, dim3(1), dim3(64), 0, st, window, 0, 0, a.W, a.H,
                           (int32_t)DSX_STRATEGY_WINDOW);
    } else if (!P.run_window) {  // the tile strategy, forced: no window
        hipLaunchKernelGGL(set_window_kernel, dim3(1), dim3(64), 0, st, window, 0, 0, 0, 0,
                           (int32_t)DSX_STRATEGY_TILES);
    } else if (d->window && d->reuse_window) {
        // the caller's stored record (its window sanitised by every kernel that reads it;
        // its strategy kept only when this call launches both)
        hipLaunchKernelGGL(copy_window_kernel, dim3(1), dim3(64), 0, st,
                           (const int32_t *)d->window, window,
                           P.run_tiles ? 0 : (int32_t)DSX_STRATEGY_WINDOW);
    } else {
        if ((s = fill_async(hist_sel, DSX_FILL_ZERO32, kSelBins, st))) return s;
        hipLaunchKernelGGL((select_hist_kernel<TC>), dim3(64), dim3(1024), 0, st, a, P.bshift,
                           P.cw, P.ch, hist_sel);
        hipLaunchKernelGGL(select_pick_kernel, dim3(1), dim3(1024), 0, st, (const uint32_t *)hist_sel,
                           P.bshift, P.cw, P.ch, a.W, a.H, (int64_t)(P.wl.cap / (uint32_t)ncat),
                           P.run_tiles ? 1 : 0, window);
        if (d->window)
            hipLaunchKernelGGL(copy_window_kernel, dim3(1), dim3(64), 0, st,
                               (const int32_t *)window, d->window, 0);
    }
    if ((s = launch_status())) return s;
    ScanArgs sa;
    sa.wl = P.wl;
    sa.window = window;
    sa.partials = ws + P.o_partials;
    sa.spill_cell = reinterpret_cast<uint32_t *>(ws + P.o_spill_cell);
    sa.spill_val = ws + P.o_spill_val;
    sa.seg_count = seg_count;
    sa.seg_cap = P.seg_cap;
    sa.hist = reinterpret_cast<uint32_t *>(ws + P.o_hist);
    sa.tile_cells = P.tl.cap;
    sa.tile_shift = P.tshift;
    sa.ntiles = P.ntiles;
    sa.o_thist = P.o_thist;
    sa.o_stage_cell = P.o_stage_cell;
    sa.o_stage_val = P.o_stage_val;
    sa.spill_nan = (P.flags & PL_NAN) ? 1u : 0u;
    sa.mean = mean;
    // the scan reads row 0 of each column even for an empty input (and drops it): give
    // an empty input a readable dummy row in the workspace
    PointArgs sa_a = a;
    if (a.n == 0) {
        sa_a.x = sa_a.y = sa_a.val = hist_sel;
        sa_a.cat.col = hist_sel;
    }
    auto scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_ANY>;
    if (a.cat.kind == DSX_CAT_NONE) scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_NONE>;
    else if ((a.cat.dtype == DSX_I8 || a.cat.dtype == DSX_U8) && a.cat.kind == DSX_CAT_CODES)
        scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_CODE8>;
    else if (a.cat.dtype == DSX_I8 || a.cat.dtype == DSX_U8)
        scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_BYTE>;
    // the common plane sets over linear axes, without categorizer: compile-time planes
    // (the linear fast path also needs the reference's own translation, t = -start * s,
    // under which in-range rows cannot project below pixel 0)
    const bool origin = a.sx > 0 && a.sy > 0 && a.tx == -(a.xmin * a.sx) &&
                        a.ty == -(a.ymin * a.sy);
    if (a.cat.kind == DSX_CAT_NONE && !a.xlog && !a.ylog && origin) {
        const uint32_t wf = P.wl.flags;
        if constexpr (!HAS_VAL) {
            if (wf == PL_CNT) scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_NONE, true, PL_CNT>;
        } else if constexpr (SQDEV) {
            if (wf == PL_SUM) scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_NONE, true, PL_SUM>;
        } else {
            if (wf == (PL_CNT | PL_SUM))
                scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_NONE, true, PL_CNT | PL_SUM>;
            else if (wf == (PL_CNTV | PL_SUM))
                scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_NONE, true, PL_CNTV | PL_SUM>;
            else if (wf == PL_CNT)
                scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_NONE, true, PL_CNT>;
        }
        // float32 columns, 16-byte aligned, whole groups of four rows: the wide loads
        const bool wide = sizeof(TC) == 4 && (!HAS_VAL || sizeof(TV) == 4) && d->n >= 4 &&
                          d->n % 4 == 0 && ((uintptr_t)a.x % 16) == 0 && ((uintptr_t)a.y % 16) == 0 &&
                          (!HAS_VAL || ((uintptr_t)a.val % 16) == 0);
        if constexpr (sizeof(TC) == 4 && (!HAS_VAL || sizeof(TV) == 4)) {
            if (wide) {
                if constexpr (!HAS_VAL) {
                    if (wf == PL_CNT) scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_NONE, true, PL_CNT, true>;
                } else if constexpr (!SQDEV) {
                    if (wf == (PL_CNT | PL_SUM))
                        scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_NONE, true, PL_CNT | PL_SUM, true>;
                    else if (wf == (PL_CNTV | PL_SUM))
                        scan = scan_kernel<TC, TV, HAS_VAL, KEYB, SQDEV, CK_NONE, true, PL_CNTV | PL_SUM, true>;
                }
            }
        }
    }
    hipError_t e;
    if (P.run_window) {
        e = hipFuncSetAttribute((const void *)scan, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)P.scan_lds);
        if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
        hipLaunchKernelGGL(scan, dim3(P.scan_groups), dim3(kScanThreads), P.scan_lds, st, sa_a, sa);
        if ((s = launch_status())) return s;
    }
    BktArgs ba;
    if (P.run_tiles) {
        ba.rec = window;
        ba.nb = P.ntiles;
        ba.tile_shift = (uint32_t)P.tshift;
        ba.cshift = P.cshift;
        ba.o_bval = P.o_bval;
        ba.blk_bytes = P.blk_bytes;
        ba.bcap = P.bcap;
        ba.pool = ws + P.o_pool;
        ba.hdr = reinterpret_cast<unsigned long long *>(ws + P.o_bhdr);
        ba.nblk = reinterpret_cast<uint32_t *>(ws + P.o_nblk);
        ba.hist = reinterpret_cast<uint32_t *>(ws + P.o_hist);
        ba.wl = reinterpret_cast<uint32_t *>(ws + P.o_bwl);
        ba.lpre = reinterpret_cast<uint32_t *>(ws + P.o_blpre);
        ba.o_cnt = P.bo_cnt;
        ba.o_lo = P.bo_lo;
        ba.o_tblk = P.bo_tblk;
        ba.o_base = P.bo_base;
        ba.o_bk = P.bo_bk;
        ba.o_fl = P.bo_fl;
        ba.o_off = P.bo_off;
        ba.o_val = P.bo_val;
        ba.o_ctrl = P.bo_ctrl;
        ba.loc_cap = P.bo_cnt / 4u;  // values and offsets come first, from byte 0
        ba.batch = P.batch;
        const bool lin = !a.xlog && !a.ylog && origin;
        auto bscan = bscan_kernel<TC, TV, HAS_VAL, CK_ANY, false>;
        if (a.cat.kind == DSX_CAT_NONE)
            bscan = lin ? bscan_kernel<TC, TV, HAS_VAL, CK_NONE, true>
                        : bscan_kernel<TC, TV, HAS_VAL, CK_NONE, false>;
        else if ((a.cat.dtype == DSX_I8 || a.cat.dtype == DSX_U8) && a.cat.kind == DSX_CAT_CODES)
            bscan = lin ? bscan_kernel<TC, TV, HAS_VAL, CK_CODE8, true>
                        : bscan_kernel<TC, TV, HAS_VAL, CK_CODE8, false>;
        // float32 columns without values, 16-byte aligned (the code column 4-byte aligned),
        // whole groups of four rows: the wide loads
        if constexpr (sizeof(TC) == 4 && !HAS_VAL) {
            const bool codes = (a.cat.dtype == DSX_I8 || a.cat.dtype == DSX_U8) &&
                               a.cat.kind == DSX_CAT_CODES;
            const bool wide = lin && d->n >= 4 && d->n % 4 == 0 && ((uintptr_t)a.x % 16) == 0 &&
                              ((uintptr_t)a.y % 16) == 0 &&
                              (a.cat.kind == DSX_CAT_NONE || (codes && ((uintptr_t)a.cat.col % 4) == 0));
            if (wide)
                bscan = a.cat.kind == DSX_CAT_NONE ? bscan_kernel<TC, TV, HAS_VAL, CK_NONE, true, true>
                                                   : bscan_kernel<TC, TV, HAS_VAL, CK_CODE8, true, true>;
        }
        else if (a.cat.dtype == DSX_I8 || a.cat.dtype == DSX_U8)
            bscan = bscan_kernel<TC, TV, HAS_VAL, CK_BYTE, false>;
        e = hipFuncSetAttribute((const void *)bscan, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)P.bscan_lds);
        if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
        hipLaunchKernelGGL(bscan, dim3(P.scan_groups), dim3(kScanThreads), P.bscan_lds, st, sa_a, ba);
        if ((s = launch_status())) return s;
    }
    if (P.run_window) {
        auto merge_w = merge_window_kernel<KEYB>;  // the common plane sets: compile-time planes
        if (P.wl.flags == PL_CNT) merge_w = merge_window_kernel<KEYB, PL_CNT>;
        else if (P.wl.flags == (PL_CNT | PL_SUM)) merge_w = merge_window_kernel<KEYB, PL_CNT | PL_SUM>;
        else if (P.wl.flags == (PL_CNTV | PL_SUM)) merge_w = merge_window_kernel<KEYB, PL_CNTV | PL_SUM>;
        hipLaunchKernelGGL(merge_w, dim3((P.wl.cap + kMergeCells - 1) / kMergeCells),
                           dim3(kMergeCells * kMergeWays), 0, st, P.wl, (const uint8_t *)sa.partials,
                           P.scan_groups, (const int32_t *)window, a.W, a.H, ncat, out);
        if ((s = launch_status())) return s;
    }
    PartArgs pa;
    pa.cell = sa.spill_cell;
    pa.val = sa.spill_val;
    pa.seg_count = seg_count;
    pa.seg_cap = P.seg_cap;
    pa.ngroups = P.scan_groups;
    pa.tile_cells = P.tl.cap;
    pa.tile_shift = P.tshift;
    pa.ntiles = P.ntiles;
    pa.hist = reinterpret_cast<uint32_t *>(ws + P.o_hist);
    pa.tile_start = reinterpret_cast<unsigned long long *>(ws + P.o_tile_start);
    pa.sorted_cell = reinterpret_cast<uint32_t *>(ws + P.o_sorted_cell);
    pa.sorted_val = ws + P.o_sorted_val;
    pa.sink = (size_t)(d->n < 1 ? 1 : d->n);
    pa.part_start = reinterpret_cast<uint32_t *>(ws + P.o_part_start);
    pa.part_len = reinterpret_cast<uint32_t *>(ws + P.o_part_len);
    pa.part_tile = reinterpret_cast<uint32_t *>(ws + P.o_part_tile);
    pa.max_parts = P.max_parts;
    pa.scratch = ws + P.o_scratch;
    pa.rec = window;
    pa.min_part_tiles = P.tiles_ok ? P.min_part_tiles : (uint32_t)kMinPart;
    unsigned long long *totals = reinterpret_cast<unsigned long long *>(ws + P.o_totals);
    hipLaunchKernelGGL(spill_colscan_kernel, dim3(P.ntiles), dim3(kPartThreads), 0, st, pa, totals);
    hipLaunchKernelGGL(spill_tilescan_kernel, dim3(1), dim3(kPartThreads), 0, st, pa,
                       (const unsigned long long *)totals);
    if (P.run_window)
        hipLaunchKernelGGL((spill_scatter_kernel<TV, HAS_VAL>), dim3(P.scan_groups),
                           dim3(kPartThreads), 0, st, pa);
    if ((s = launch_status())) return s;
    if (P.run_window) {
        auto tile = tile_agg_kernel<TV, HAS_VAL, KEYB, SQDEV>;
        if constexpr (!SQDEV) {  // the common plane sets: compile-time planes
            if (P.tl.flags == PL_CNT) tile = tile_agg_kernel<TV, HAS_VAL, KEYB, SQDEV, PL_CNT>;
            else if (P.tl.flags == (PL_CNT | PL_SUM))
                tile = tile_agg_kernel<TV, HAS_VAL, KEYB, SQDEV, PL_CNT | PL_SUM>;
            else if (P.tl.flags == (PL_CNTV | PL_SUM))
                tile = tile_agg_kernel<TV, HAS_VAL, KEYB, SQDEV, PL_CNTV | PL_SUM>;
            else if (P.tl.flags == (PL_CNT | PL_SUM | PL_NAN))  // count + mean / sum
                tile = tile_agg_kernel<TV, HAS_VAL, KEYB, SQDEV, PL_CNT | PL_SUM | PL_NAN>;
        }
        e = hipFuncSetAttribute((const void *)tile, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)P.tile_lds);
        if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
        hipLaunchKernelGGL(tile, dim3(P.max_parts), dim3(kTileThreads), P.tile_lds, st, pa, P.tl,
                           (const int32_t *)window, a.W, a.H, wcap_px, ncat, P.ncells, mean, out);
        if ((s = launch_status())) return s;
    }
    if (P.run_tiles) {
        auto btile = btile_agg_kernel<TV, HAS_VAL, KEYB, SQDEV>;
        e = hipFuncSetAttribute((const void *)btile, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)P.btile_lds);
        if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
        hipLaunchKernelGGL(btile, dim3(P.max_parts), dim3(kTileThreads), P.btile_lds, st, pa, ba,
                           P.tl, (const int32_t *)window, a.W, a.H, wcap_px, ncat, P.ncells, mean,
                           out);
        if ((s = launch_status())) return s;
    }
    auto merge = tile_merge_kernel<KEYB>;
    e = hipFuncSetAttribute((const void *)merge, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)P.tile_lds);
    if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
    hipLaunchKernelGGL(merge, dim3(P.ntiles), dim3(kTileThreads), P.tile_lds, st, pa, P.tl,
                       (const int32_t *)window, a.W, a.H, wcap_px, ncat, P.ncells, out);
    return launch_status();
}

// ---- canvases beyond the tiles: one pass of global atomics ------------------------------
// Every kept row updates its cell of the output planes directly (the planes start at the
// reference's initial values), as the reference's append does (reductions.py).
template <typename TC, typename TV, bool HAS_VAL, bool SQDEV>
__global__ __launch_bounds__(256) void points_global_kernel(PointArgs a, Planes out,
                                                            const double *mean) {
    const TC *X = reinterpret_cast<const TC *>(a.x);
    const TC *Y = reinterpret_cast<const TC *>(a.y);
    const TV *V = reinterpret_cast<const TV *>(a.val);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        int32_t px, py, pc;
        if (!locate<TC>(a, i, stream_load(X + i), stream_load(Y + i), px, py, pc)) continue;
        const int64_t cell = ((int64_t)py * a.W + px) * a.cat.ncat + pc;
        TV v = TV(0);
        if constexpr (HAS_VAL) v = stream_load(V + i);
        if constexpr (SQDEV) {
            if (v != v) continue;
            const double dd = (double)v - mean[cell];
            unsafeAtomicAdd(out.sum + cell, dd * dd);
        } else {
            append<TV, HAS_VAL>(out, cell, v);
        }
    }
}

template <typename TC>
static int run_global(const dsx_points_desc *d, const PointArgs &a, const dsx_planes *pl_in,
                      const Planes &out_in, const double *mean, uint8_t *ws, hipStream_t s) {
    const int64_t ncells = (int64_t)a.W * a.H * a.cat.ncat;
    const bool f32 = has_value(d) && d->val_dtype == DSX_F32;
    int st;
    // a finalized mean: its sum and count(v) planes (the caller's, or the workspace's)
    dsx_planes plv = *pl_in;
    Planes out = out_in;
    if (pl_in->mean) {
        size_t o = 256;
        if (!plv.sum) { plv.sum = reinterpret_cast<double *>(ws + o); o += align_up(8 * (size_t)ncells, 256); }
        if (!plv.countv) plv.countv = reinterpret_cast<uint32_t *>(ws + o);
        out.sum = plv.sum;
        out.countv = plv.countv;
        out.mean = nullptr;  // finalized after the pass
    }
    const dsx_planes *pl = &plv;
    if (pl->count && (st = fill_async(pl->count, DSX_FILL_ZERO32, ncells, s))) return st;
    if (pl->countv && (st = fill_async(pl->countv, DSX_FILL_ZERO32, ncells, s))) return st;
    if (pl->sum && (st = fill_async(pl->sum, DSX_FILL_ZERO64, ncells, s))) return st;
    if (pl->min && (st = fill_async(pl->min, f32 ? DSX_FILL_ONES32 : DSX_FILL_ONES64, ncells, s)))
        return st;
    if (pl->max && (st = fill_async(pl->max, f32 ? DSX_FILL_ZERO32 : DSX_FILL_ZERO64, ncells, s)))
        return st;
    const int grid = grid_for(d->n < 1 ? 1 : d->n, 256, 8);
    if (d->n == 0) {
        // no rows: the filled planes are the result (a mean is finalized below)
    } else if (!has_value(d))
        hipLaunchKernelGGL((points_global_kernel<TC, float, false, false>), dim3(grid), dim3(256), 0, s, a, out, mean);
    else if (f32 && mean)
        hipLaunchKernelGGL((points_global_kernel<TC, float, true, true>), dim3(grid), dim3(256), 0, s, a, out, mean);
    else if (f32)
        hipLaunchKernelGGL((points_global_kernel<TC, float, true, false>), dim3(grid), dim3(256), 0, s, a, out, mean);
    else if (mean)
        hipLaunchKernelGGL((points_global_kernel<TC, double, true, true>), dim3(grid), dim3(256), 0, s, a, out, mean);
    else
        hipLaunchKernelGGL((points_global_kernel<TC, double, true, false>), dim3(grid), dim3(256), 0, s, a, out, mean);
    if ((st = launch_status())) return st;
    if (pl_in->mean) {
        dsx_finalize_desc f;
        memset(&f, 0, sizeof f);
        f.mode = DSX_FIN_MEAN;
        f.ncells = ncells;
        f.countv = plv.countv;
        f.sum = plv.sum;
        f.out = pl_in->mean;
        return dsx_finalize(&f, s);
    }
    return DSX_OK;
}

template <typename TC>
static int dispatch_val(const dsx_points_desc *d, const PointArgs &a, const Plan &P,
                        const Planes &out, const double *mean, uint8_t *ws, hipStream_t st) {
    const bool sq = mean != nullptr;
    if (!has_value(d)) return run_points<TC, float, false, 8, false>(d, a, P, out, mean, ws, st);
    if (d->val_dtype == DSX_F32)
        return sq ? run_points<TC, float, true, 4, true>(d, a, P, out, mean, ws, st)
                  : run_points<TC, float, true, 4, false>(d, a, P, out, mean, ws, st);
    return sq ? run_points<TC, double, true, 8, true>(d, a, P, out, mean, ws, st)
              : run_points<TC, double, true, 8, false>(d, a, P, out, mean, ws, st);
}

// ---- indexed aggregation (rows sorted by the cells of a lattice) ------------------------
// The canvas is cut into tiles of tw x th pixels whose planes fit LDS next to a table of
// the lattice rows under the tile.  Tile t reads the rows of the lattice cells its pixels
// can receive (one extra cell on each side), keeps the rows whose pixel is in the tile
// and writes every cell of the tile: each row lands in exactly one tile, and no row is
// spilled.  A tile with many rows is split into parts of at most part_len rows, whose
// regions a merge kernel folds, as for the spill tiles.
constexpr int kIdxThreads = 1024;
constexpr int kIdxMaxRows = 1024;       // lattice rows a tile may span (grid <= 1024)
constexpr int kIdxTable = kIdxMaxRows * 16 + 16;
constexpr unsigned long long kIdxMinPart = 1ull << 16;

struct IdxArgs {
    PointArgs a;                     // the sorted columns and the view
    Lay L;                           // tile planes: cap = tw * th cells
    int32_t tw, th, ntx, nty, ntiles;
    int32_t grid;
    double ex0, esx, ey0, esy;       // lattice: cell = (v - e0) * es
    const int64_t *cs;               // cell_start
    unsigned long long *tile_rows;   // [ntiles]
    uint32_t *part_start;            // [ntiles + 1]
    uint32_t *part_tile;             // [max_parts]
    unsigned long long *part_len;    // [1]
    int32_t max_parts, cus;
    uint8_t *scratch;
    const double *mean;
    Planes out;
};

__device__ __forceinline__ int idx_cell(double v, double e0, double es, int grid) {
    const double top = (double)(grid - 1);
    double f = (v - e0) * es;
    f = f < 0.0 ? 0.0 : (f > top ? top : f);
    return (int)f;
}

// Pixels and lattice cells of tile t.
__device__ __forceinline__ void idx_rect(const IdxArgs &A, int t, int &px0, int &px1, int &py0,
                                         int &py1, int &gx0, int &gx1, int &gy0, int &gy1) {
    const PointArgs &a = A.a;
    const int tx = t % A.ntx, ty = t / A.ntx;
    px0 = tx * A.tw;
    px1 = px0 + A.tw < a.W ? px0 + A.tw : a.W;
    py0 = ty * A.th;
    py1 = py0 + A.th < a.H ? py0 + A.th : a.H;
    // the coordinates that project into the pixels (sx, sy > 0); the last column / row
    // also receives the rows clamped from W / H, i.e. up to the view's upper bound
    double xl = ((double)px0 - a.tx) / a.sx, xh = ((double)px1 - a.tx) / a.sx;
    double yl = ((double)py0 - a.ty) / a.sy, yh = ((double)py1 - a.ty) / a.sy;
    xl = xl > a.xmin ? xl : a.xmin;
    yl = yl > a.ymin ? yl : a.ymin;
    xh = (px1 == a.W || xh > a.xmax) ? a.xmax : xh;
    yh = (py1 == a.H || yh > a.ymax) ? a.ymax : yh;
    gx0 = idx_cell(xl, A.ex0, A.esx, A.grid) - 1;
    gx1 = idx_cell(xh, A.ex0, A.esx, A.grid) + 1;
    gy0 = idx_cell(yl, A.ey0, A.esy, A.grid) - 1;
    gy1 = idx_cell(yh, A.ey0, A.esy, A.grid) + 1;
    gx0 = gx0 < 0 ? 0 : gx0;
    gy0 = gy0 < 0 ? 0 : gy0;
    gx1 = gx1 >= A.grid ? A.grid - 1 : gx1;
    gy1 = gy1 >= A.grid ? A.grid - 1 : gy1;
}

// Rows under each tile (one workgroup per tile).
__global__ __launch_bounds__(256) void idx_rows_kernel(IdxArgs A) {
    __shared__ unsigned long long red[256 / 64];
    const int t = blockIdx.x;
    int px0, px1, py0, py1, gx0, gx1, gy0, gy1;
    idx_rect(A, t, px0, px1, py0, py1, gx0, gx1, gy0, gy1);
    unsigned long long sum = 0;
    for (int gy = gy0 + threadIdx.x; gy <= gy1; gy += blockDim.x) {
        const int64_t r = (int64_t)gy * A.grid;
        sum += (unsigned long long)(A.cs[r + gx1 + 1] - A.cs[r + gx0]);
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s2 = 0;
        for (int k = 0; k < 256 / 64; ++k) s2 += red[k];
        A.tile_rows[t] = s2;
    }
}

// Parts of at most E rows, E = max(kIdxMinPart, total / (2 * CUs)): at most
// ntiles + 2 * CUs parts.
__global__ __launch_bounds__(kPartThreads) void idx_plan_kernel(IdxArgs A) {
    __shared__ unsigned long long sh[kPartThreads / 64 + 1];
    unsigned long long total = 0;
    for (int t0 = 0; t0 < A.ntiles; t0 += blockDim.x) {
        const int t = t0 + threadIdx.x;
        const unsigned long long v = t < A.ntiles ? A.tile_rows[t] : 0ull;
        unsigned long long tot;
        block_exscan(v, sh, tot);
        total += tot;
    }
    const unsigned long long share = (total + 2ull * A.cus - 1) / (2ull * A.cus);
    const unsigned long long E = share > kIdxMinPart ? share : kIdxMinPart;
    if (threadIdx.x == 0) *A.part_len = E;
    unsigned long long pcarry = 0;
    for (int t0 = 0; t0 < A.ntiles; t0 += blockDim.x) {
        const int t = t0 + threadIdx.x;
        unsigned long long np = 0;
        if (t < A.ntiles) {
            np = (A.tile_rows[t] + E - 1) / E;
            np = np < 1 ? 1 : np;
        }
        unsigned long long tot;
        const unsigned long long ex = block_exscan(np, sh, tot);
        if (t < A.ntiles) {
            A.part_start[t] = (uint32_t)(pcarry + ex);
            for (unsigned long long q = 0; q < np; ++q)
                if (pcarry + ex + q < (unsigned long long)A.max_parts)  // always (bound above)
                    A.part_tile[pcarry + ex + q] = (uint32_t)t;
        }
        pcarry += tot;
    }
    if (threadIdx.x == 0) A.part_start[A.ntiles] = (uint32_t)pcarry;
}

// Write an indexed tile's pixels [px0, px1) x [py0, py1) from its LDS region.
template <int KEYB>
__device__ void idx_store(const uint8_t *smem, const IdxArgs &A, int px0, int px1, int py0,
                          int py1) {
    const int w = px1 - px0, h = py1 - py0;
    for (int c = threadIdx.x; c < w * h; c += kIdxThreads) {
        const int ly = c / w, lx = c - ly * w;
        region_store<KEYB>(smem, A.L, (uint32_t)(ly * A.tw + lx), A.out,
                           (int64_t)(py0 + ly) * A.a.W + (px0 + lx));
    }
}

template <typename TC, typename TV, bool HAS_VAL, int KEYB, bool SQDEV>
__global__ __launch_bounds__(kIdxThreads) void idx_tile_kernel(IdxArgs A) {
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ unsigned long long sh[kIdxThreads / 64 + 1];
    const PointArgs &a = A.a;
    const Lay &L = A.L;
    const uint32_t part = blockIdx.x;
    if (part >= A.part_start[A.ntiles] || part >= (uint32_t)A.max_parts) return;
    const int t = (int)A.part_tile[part];
    const uint32_t first = A.part_start[t], nparts = A.part_start[t + 1] - first;
    int px0, px1, py0, py1, gx0, gx1, gy0, gy1;
    idx_rect(A, t, px0, px1, py0, py1, gx0, gx1, gy0, gy1);
    const int nrow = gy1 - gy0 + 1;  // <= kIdxMaxRows (grid <= 1024)
    // table of the lattice rows under the tile: first sorted row and row-count prefix
    int64_t *s_start = reinterpret_cast<int64_t *>(smem + L.bytes);
    unsigned long long *s_pre = reinterpret_cast<unsigned long long *>(smem + L.bytes) + kIdxMaxRows;
    unsigned long long len = 0;
    int64_t st0 = 0;
    if ((int)threadIdx.x < nrow) {
        const int64_t r = (int64_t)(gy0 + (int)threadIdx.x) * A.grid;
        st0 = A.cs[r + gx0];
        len = (unsigned long long)(A.cs[r + gx1 + 1] - st0);
    }
    unsigned long long tot;
    const unsigned long long ex = block_exscan(len, sh, tot);
    if ((int)threadIdx.x < nrow) {
        s_start[threadIdx.x] = st0;
        s_pre[threadIdx.x] = ex;
    }
    region_init<KEYB>(smem, L, threadIdx.x, kIdxThreads);
    __syncthreads();
    const unsigned long long E = *A.part_len;
    unsigned long long lo = (unsigned long long)(part - first) * E, hi = lo + E;
    lo = lo < tot ? lo : tot;
    hi = hi < tot ? hi : tot;
    const TC *X = reinterpret_cast<const TC *>(a.x);
    const TC *Y = reinterpret_cast<const TC *>(a.y);
    const TV *V = reinterpret_cast<const TV *>(a.val);
    for (unsigned long long c0 = lo; c0 < hi; c0 += (unsigned long long)kChunk) {
        TC xs[kItems], ys[kItems];
        TV vs[kItems];
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            unsigned long long j = c0 + (unsigned long long)(it * kIdxThreads + threadIdx.x);
            j = j < hi ? j : lo;
            int k = 0;  // the last lattice row starting at or before j
            for (int step = kIdxMaxRows / 2; step >= 1; step >>= 1)
                if (k + step < nrow && s_pre[k + step] <= j) k += step;
            const int64_t row = s_start[k] + (int64_t)(j - s_pre[k]);
            xs[it] = X[row];
            ys[it] = Y[row];
            if constexpr (HAS_VAL) vs[it] = V[row];
            else vs[it] = TV(0);
        }
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            const unsigned long long j = c0 + (unsigned long long)(it * kIdxThreads + threadIdx.x);
            int32_t px, py;
            if (j >= hi || !locate_xy<TC>(a, xs[it], ys[it], px, py)) continue;
            if (px < px0 || px >= px1 || py < py0 || py >= py1) continue;  // another tile's
            const uint32_t lc = (uint32_t)((py - py0) * A.tw + (px - px0));
            const double v = HAS_VAL ? (double)vs[it] : 0.0;
            const bool vnan = HAS_VAL ? (v != v) : true;
            double add = v;
            if constexpr (SQDEV) {
                const double d = v - A.mean[(int64_t)py * a.W + px];
                add = d * d;
            }
            region_add<KEYB>(smem, L, lc, vnan, v, add);
        }
    }
    __syncthreads();
    if (nparts > 1) {
        // a split tile: each part leaves its region in scratch; idx_merge_kernel,
        // launched next, folds them and writes the tile (no device-scope fence per part)
        uint4 *mine = reinterpret_cast<uint4 *>(A.scratch + (size_t)part * L.bytes);
        const uint4 *src = reinterpret_cast<const uint4 *>(smem);
        for (uint32_t j = threadIdx.x; j < L.bytes / 16u; j += kIdxThreads) mine[j] = src[j];
        return;
    }
    idx_store<KEYB>(smem, A, px0, px1, py0, py1);
}

// Split indexed tiles: fold the parts' regions, write the tile.
template <int KEYB>
__global__ __launch_bounds__(kIdxThreads) void idx_merge_kernel(IdxArgs A) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int t = blockIdx.x;
    if (t >= A.ntiles) return;
    const uint32_t first = A.part_start[t], nparts = A.part_start[t + 1] - first;
    if (nparts <= 1) return;  // uniform
    int px0, px1, py0, py1, gx0, gx1, gy0, gy1;
    idx_rect(A, t, px0, px1, py0, py1, gx0, gx1, gy0, gy1);
    fold_parts<kIdxThreads, KEYB>(smem, A.L, A.scratch + (size_t)first * A.L.bytes, nparts);
    __syncthreads();
    idx_store<KEYB>(smem, A, px0, px1, py0, py1);
}

struct IdxPlan {
    Lay L;
    int32_t tw, th, ntx, nty, ntiles, max_parts, cus;
    size_t lds, o_rows, o_pstart, o_ptile, o_plen, o_scratch, total;
};

static int idx_make_plan(const dsx_points_desc *d, const dsx_view *v, const dsx_planes *pl,
                         IdxPlan &P) {
    if (plane_flags(pl) == 0) return DSX_ERR_ACCUM;
    const uint32_t flags = internal_flags(plane_flags(pl));
    const int keyb = (has_value(d) && d->val_dtype == DSX_F32) ? 4 : 8;
    const uint32_t cb = cell_bytes(flags, keyb);
    // LDS: planes + lattice-row table + the kernel's static words (under 512 bytes)
    const int64_t cap = (int64_t)(kLdsBytes - kIdxTable - 512) / cb;
    int th = (int)sqrt((double)cap);
    th = th < 1 ? 1 : (th > v->height ? v->height : th);
    int tw = (int)(cap / th);
    tw = tw > v->width ? v->width : tw;
    if (tw < 1) return DSX_ERR_SHAPE;
    P.tw = tw; P.th = th;
    P.ntx = (v->width + tw - 1) / tw;
    P.nty = (v->height + th - 1) / th;
    if ((int64_t)P.ntx * P.nty > (1 << 24)) return DSX_ERR_SHAPE;
    P.ntiles = P.ntx * P.nty;
    P.L = make_lay(flags, keyb, (uint32_t)(tw * th));
    P.lds = P.L.bytes + kIdxTable;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && c > 0)
            cus = c;
    }
    P.cus = cus;
    P.max_parts = P.ntiles + 2 * cus + 1;
    size_t o = 0;
    P.o_rows = o; o = align_up(o + 8 * (size_t)P.ntiles, 256);
    P.o_pstart = o; o = align_up(o + 4 * (size_t)(P.ntiles + 1), 256);
    P.o_ptile = o; o = align_up(o + 4 * (size_t)P.max_parts, 256);
    P.o_plen = o; o = align_up(o + 8, 256);
    P.o_scratch = o; o = align_up(o + (size_t)P.max_parts * P.L.bytes, 256);
    P.total = o;
    return DSX_OK;
}

static int idx_check(const dsx_points_desc *d, const dsx_view *v, const dsx_planes *p,
                     const double *mean, const dsx_grid_ref *g) {
    int st = check_points(d, v);
    if (st) return st;
    if (!p || !g) return DSX_ERR_NULL;
    if (d->cat.kind != DSX_CAT_NONE || v->xlog || v->ylog) return DSX_ERR_MODE;
    if (!(v->sx > 0) || !(v->sy > 0)) return DSX_ERR_SHAPE;
    if (g->grid < 1 || g->grid > kIdxMaxRows) return DSX_ERR_SHAPE;
    if (!(g->x1 > g->x0) || !(g->y1 > g->y0)) return DSX_ERR_SHAPE;
    if (!g->cell_start) return DSX_ERR_NULL;
    const bool needs_val = p->countv || p->sum || p->min || p->max || p->mean;
    if (needs_val && !has_value(d)) return DSX_ERR_ACCUM;
    if (mean && (!has_value(d) || !p->sum || p->count || p->countv || p->min || p->max || p->mean))
        return DSX_ERR_ACCUM;
    return DSX_OK;
}

template <typename TC, typename TV, bool HAS_VAL, int KEYB, bool SQDEV>
static int idx_run(IdxArgs &A, const IdxPlan &P, hipStream_t st) {
    hipLaunchKernelGGL(idx_rows_kernel, dim3(P.ntiles), dim3(256), 0, st, A);
    hipLaunchKernelGGL(idx_plan_kernel, dim3(1), dim3(kPartThreads), 0, st, A);
    int s;
    if ((s = launch_status())) return s;
    auto kern = idx_tile_kernel<TC, TV, HAS_VAL, KEYB, SQDEV>;
    const hipError_t e = hipFuncSetAttribute((const void *)kern,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)P.lds);
    if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
    hipLaunchKernelGGL(kern, dim3(P.max_parts), dim3(kIdxThreads), P.lds, st, A);
    if ((s = launch_status())) return s;
    auto merge = idx_merge_kernel<KEYB>;
    const hipError_t e2 = hipFuncSetAttribute((const void *)merge,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)A.L.bytes);
    if (e2 != hipSuccess) { record_hip_error(e2); return DSX_ERR_HIP; }
    hipLaunchKernelGGL(merge, dim3(P.ntiles), dim3(kIdxThreads), A.L.bytes, st, A);
    return launch_status();
}

template <typename TC>
static int idx_dispatch(const dsx_points_desc *d, IdxArgs &A, const IdxPlan &P, hipStream_t st) {
    const bool sq = A.mean != nullptr;
    if (!has_value(d)) return idx_run<TC, float, false, 8, false>(A, P, st);
    if (d->val_dtype == DSX_F32)
        return sq ? idx_run<TC, float, true, 4, true>(A, P, st)
                  : idx_run<TC, float, true, 4, false>(A, P, st);
    return sq ? idx_run<TC, double, true, 8, true>(A, P, st)
              : idx_run<TC, double, true, 8, false>(A, P, st);
}

}  // namespace dsx

using namespace dsx;

extern "C" size_t dsx_points_workspace_size(const dsx_points_desc *d, const dsx_view *v,
                                            const dsx_planes *p) {
    if (check_points(d, v) || !p) return 0;
    Plan P;
    if (make_plan(d, v, p, P)) return 0;
    return P.total;
}

extern "C" int dsx_points_agg(const dsx_points_desc *d, const dsx_view *v, const dsx_planes *p,
                              const double *mean, void *workspace, size_t workspace_bytes,
                              void *stream) {
    int st = check_points(d, v);
    if (st) return st;
    if (!p) return DSX_ERR_NULL;
    const bool needs_val = p->countv || p->sum || p->min || p->max || p->mean;
    if (needs_val && !has_value(d)) return DSX_ERR_ACCUM;
    if (mean && (!has_value(d) || !p->sum || p->count || p->countv || p->min || p->max || p->mean))
        return DSX_ERR_ACCUM;
    if (p->mean && !has_value(d)) return DSX_ERR_ACCUM;
    Plan P;
    if ((st = make_plan(d, v, p, P))) return st;
    if (!workspace) return DSX_ERR_WORKSPACE;
    if (workspace_bytes < P.total) return DSX_ERR_WORKSPACE;
    const PointArgs a = args_of(d, v);
    const Planes out = planes_of(p);
    hipStream_t s = as_stream(stream);
    uint8_t *ws = reinterpret_cast<uint8_t *>(workspace);
    if (P.global) {
        if (d->xy_dtype == DSX_F32) return run_global<float>(d, a, p, out, mean, ws, s);
        return run_global<double>(d, a, p, out, mean, ws, s);
    }
    if (d->xy_dtype == DSX_F32) return dispatch_val<float>(d, a, P, out, mean, ws, s);
    return dispatch_val<double>(d, a, P, out, mean, ws, s);
}

// ---- row-index reductions ---------------------------------------------------------------
// One pass over the rows with a global atomic per candidate row; a plain read of the cell
// first skips rows that cannot improve it (FIRST / ARG* see rows in increasing order,
// LAST in decreasing order, so after the first sweep almost every row is skipped).

struct RiArgs {
    PointArgs a;
    int32_t mode, nan_check;
    const void *key;
    unsigned long long *index;
};

// Each thread takes kRiU rows per iteration (a grid stride apart) and issues the loads
// of all of them before it waits for any: the argmax / argmin pass is bound by the
// latency of its dependent reads (column -> cell -> key plane -> index plane), so rows in
// flight is what counts (where(max) 1.36 -> 1.06 ms on 100M rows).  first / last keep
// one row per iteration: their skip test works best with the fewest rows in flight
// (the rows in flight before a cell is first set all take the atomic; 4 rows per
// iteration measured 1.15 -> 1.60 ms).  The result does not depend on the order rows
// are visited in (atomic min / max).
template <typename TC, typename TV, int kRiU>
__global__ __launch_bounds__(256) void rowindex_kernel(RiArgs r) {
    const PointArgs &a = r.a;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const bool last = r.mode == DSX_RI_LAST;
    const bool arg = r.mode == DSX_RI_ARGMAX || r.mode == DSX_RI_ARGMIN;
    for (int64_t k0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k0 < a.n;
         k0 += kRiU * stride) {
        int64_t row[kRiU];
        bool ok[kRiU];
        TC xs[kRiU], ys[kRiU];
        TV vs[kRiU];
#pragma unroll
        for (int u = 0; u < kRiU; ++u) {  // row k0 is in range, so rows past n re-read it
            const int64_t k = k0 + u * stride;
            ok[u] = k < a.n;
            const int64_t kk = ok[u] ? k : k0;
            row[u] = last ? a.n - 1 - kk : kk;
            xs[u] = stream_load(reinterpret_cast<const TC *>(a.x) + row[u]);
            ys[u] = stream_load(reinterpret_cast<const TC *>(a.y) + row[u]);
            vs[u] = r.nan_check ? stream_load(reinterpret_cast<const TV *>(a.val) + row[u]) : TV(0);
        }
        int64_t cell[kRiU];
#pragma unroll
        for (int u = 0; u < kRiU; ++u) {
            int32_t px = 0, py = 0, pc = 0;
            ok[u] = ok[u] && locate<TC>(a, row[u], xs[u], ys[u], px, py, pc);
            cell[u] = ok[u] ? ((int64_t)py * a.W + px) * a.cat.ncat + pc : 0;
            if (r.nan_check) ok[u] = ok[u] && !(vs[u] != vs[u]);
        }
        if (arg) {  // keep rows whose value is the cell's extreme
#pragma unroll
            for (int u = 0; u < kRiU; ++u) {
                if (!ok[u]) continue;
                if constexpr (sizeof(TV) == 4) {
                    ok[u] = fkey32((float)vs[u]) == reinterpret_cast<const uint32_t *>(r.key)[cell[u]];
                } else {
                    ok[u] = fkey64((double)vs[u]) == reinterpret_cast<const uint64_t *>(r.key)[cell[u]];
                }
            }
        }
        long long cur[kRiU];
#pragma unroll
        // a relaxed device-scope read first (served by L2): most rows cannot improve the cell
        for (int u = 0; u < kRiU; ++u)
            cur[u] = ok[u] ? (long long)__hip_atomic_load(r.index + cell[u], __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT)
                           : 0;
#pragma unroll
        for (int u = 0; u < kRiU; ++u) {
            if (!ok[u]) continue;
            unsigned long long *slot = r.index + cell[u];
            if (last) {
                if ((long long)row[u] > cur[u]) atomicMax(reinterpret_cast<long long *>(slot), (long long)row[u]);
            } else {
                // -1 (empty) is the largest unsigned value, so min keeps the smallest row
                if ((unsigned long long)row[u] < (unsigned long long)cur[u])
                    atomicMin(slot, (unsigned long long)row[u]);
            }
        }
    }
}

// where(max / min): the first row whose value equals its cell's extreme (the key plane of
// the max / min pass), as one streaming pass like the points scan (chunks of 4 096 rows,
// three rotating register sets, the linear projection and byte-code categories
// specialised).  Each kept row reads its cell's key (L2-resident plane); the reads are
// issued before the loads of the chunk two ahead and waited for after them, so the wait
// does not also wait for those loads.  A row matching the key takes the global atomic
// min (rows matching an extreme are few: no skip read).
template <typename TC, typename TV, int CK, bool LIN>
__global__ __launch_bounds__(kScanThreads) void rowindex_arg_kernel(RiArgs r) {
    const PointArgs &a = r.a;
    const int tid = threadIdx.x;
    const int ncat = a.cat.ncat;
    using K = KeyOf<sizeof(TV)>;
    const auto *const key = global_ptr(reinterpret_cast<const typename K::T *>(r.key));
    auto *const index = global_ptr(r.index);
    const int64_t nchunks = (a.n + kChunk - 1) / kChunk;
    using R = Rows<TC, TV, true, CK>;
    auto step = [&](int64_t ck, const R &S, R &far) {
        const int64_t b = ck * kChunk;
        const int64_t rem64 = a.n - b;
        const uint32_t rem = rem64 >= kChunk ? (uint32_t)kChunk : (rem64 > 0 ? (uint32_t)rem64 : 0u);
        int64_t cell[kItems];
        bool keep[kItems];
        typename K::T kv[kItems];
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            const uint32_t j = (uint32_t)(it * kScanThreads + tid);
            int32_t px = 0, py = 0, pc = 0;
            bool k_;
            if constexpr (LIN) {  // as in scan_kernel
                const double x = (double)S.x[it], y = (double)S.y[it];
                const bool inr = (a.xmin <= x) & (x <= a.xmax) & (a.ymin <= y) & (y <= a.ymax);
                const double fx = inr ? __dadd_rn(__dmul_rn(x, a.sx), a.tx) : 0.0;
                const double fy = inr ? __dadd_rn(__dmul_rn(y, a.sy), a.ty) : 0.0;
                const int32_t xx = (int32_t)fx, yy = (int32_t)fy;
                px = xx < a.W - 1 ? xx : a.W - 1;
                py = yy < a.H - 1 ? yy : a.H - 1;
                k_ = (j < rem) & inr;
            } else {
                k_ = j < rem && locate_xy<TC>(a, S.x[it], S.y[it], px, py);
            }
            if constexpr (CK != CK_NONE) {
                pc = scan_category<CK>(a.cat, S.c[it]);
                k_ = k_ && pc >= 0;
            }
            k_ = k_ && !(S.v[it] != S.v[it]);
            keep[it] = k_;
            cell[it] = k_ ? ((int64_t)py * a.W + px) * ncat + pc : 0;
        }
#pragma unroll
        for (int it = 0; it < kItems; ++it) kv[it] = keep[it] ? key[cell[it]] : (typename K::T)0;
        far.load(a, ck + 2 * (int64_t)gridDim.x, tid);
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            if (!keep[it] || K::key((double)S.v[it]) != kv[it]) continue;
            const unsigned long long row = (unsigned long long)(b + it * kScanThreads + tid);
            __hip_atomic_fetch_min(index + cell[it], row, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    R sa, sb, sc;
    const int64_t G = gridDim.x;
    sa.load(a, blockIdx.x, tid);
    sb.load(a, blockIdx.x + G, tid);
    for (int64_t ck = blockIdx.x; ck < nchunks; ck += 3 * G) {  // steps past the end drop rows
        step(ck, sa, sc);
        step(ck + G, sb, sa);
        step(ck + 2 * G, sc, sb);
    }
}

template <typename TC>
static int rowindex_launch(const RiArgs &r, int val_dtype, hipStream_t st) {
    const bool arg = r.mode == DSX_RI_ARGMAX || r.mode == DSX_RI_ARGMIN;
    if (arg) {  // the streaming form
        const PointArgs &a = r.a;
        const bool lin = !a.xlog && !a.ylog && a.sx > 0 && a.sy > 0 &&
                         a.tx == -(a.xmin * a.sx) && a.ty == -(a.ymin * a.sy);
        const bool byte = (a.cat.dtype == DSX_I8 || a.cat.dtype == DSX_U8) &&
                          a.cat.kind == DSX_CAT_CODES;
        const int64_t chunks = (a.n + kChunk - 1) / kChunk;
        const int groups = grid_for(chunks < 1 ? 1 : chunks, 1, 1);  // one per CU at most
        auto kern = rowindex_arg_kernel<TC, double, CK_ANY, false>;
        if (val_dtype == DSX_F32) {
            kern = rowindex_arg_kernel<TC, float, CK_ANY, false>;
            if (a.cat.kind == DSX_CAT_NONE) kern = lin ? rowindex_arg_kernel<TC, float, CK_NONE, true>
                                                       : rowindex_arg_kernel<TC, float, CK_NONE, false>;
            else if (byte) kern = rowindex_arg_kernel<TC, float, CK_CODE8, false>;
        } else {
            if (a.cat.kind == DSX_CAT_NONE) kern = lin ? rowindex_arg_kernel<TC, double, CK_NONE, true>
                                                       : rowindex_arg_kernel<TC, double, CK_NONE, false>;
            else if (byte) kern = rowindex_arg_kernel<TC, double, CK_CODE8, false>;
        }
        hipLaunchKernelGGL(kern, dim3(groups), dim3(kScanThreads), 0, st, r);
        return launch_status();
    }
    // four workgroups per CU: fewer rows in flight for the skip test (on 100M rows
    // first 0.97 -> 0.90 ms, where(max) 1.05 -> 1.01 ms against eight; two is slower)
    const int grid = grid_for(r.a.n < 1 ? 1 : r.a.n, 256, 4);
    if (val_dtype == DSX_F32) {
        if (arg) hipLaunchKernelGGL((rowindex_kernel<TC, float, 4>), dim3(grid), dim3(256), 0, st, r);
        else hipLaunchKernelGGL((rowindex_kernel<TC, float, 1>), dim3(grid), dim3(256), 0, st, r);
    } else {
        if (arg) hipLaunchKernelGGL((rowindex_kernel<TC, double, 4>), dim3(grid), dim3(256), 0, st, r);
        else hipLaunchKernelGGL((rowindex_kernel<TC, double, 1>), dim3(grid), dim3(256), 0, st, r);
    }
    return launch_status();
}

// ---- row-index reductions with an LDS window ---------------------------------------------
// FIRST / LAST: the rows that land in the call's hot window (picked from a sample of the
// rows, as for dsx_points_agg, but holding one 32-bit row per cell, so four times the
// cells) keep their cell's smallest / largest row in a per-workgroup LDS copy of the
// window (an LDS atomic min / max); the other rows take the global pass's atomic with its
// skip test.  A merge then folds the workgroups' windows into the index plane (the global
// atomics never touch a window cell).  Rows are < 2^32 (checked on the host), so a row fits
// 32 bits: LAST keeps row + 1, so that 0 is the empty cell.
// ARGMAX / ARGMIN over a float32 value column without a key plane (PK): one such pass over
// 64-bit cells holding key(v) << 32 | row (ARGMIN, atomic min) or key(v) << 32 | ~row
// (ARGMAX, atomic max): the extreme value wins and, among its rows, the smallest -- the
// row the reference keeps (its strict > / < appends, reductions.py:1160-1260 and
// 2236-2372) -- so the max / min pass and the key-matching pass become one.
constexpr int kRiwThreads = 1024;
constexpr int kRiwItems = 4;
constexpr int kRiwChunk = kRiwThreads * kRiwItems;

struct RiwArgs {
    PointArgs a;
    int32_t last, nan_check;      // (packed: last = the largest packed key wins)
    unsigned long long *index;    // cell c's slot: index[c * nsel + kslot]
    const int32_t *window;
    uint32_t cap;                 // window cells (pixels x categories)
    uint32_t *partials;           // [G][cap] (packed: 64-bit cells)
    int32_t packed;
    int32_t pkmode;               // packed key: 0 key(v) << 32 | row (~row when last),
                                  // 1 row, 2 ~row (32-bit)
    int32_t nsel, kslot;
    int32_t thr;                  // packed: keep only keys past the previous slot's (kslot > 0)
};

struct RiwPlan {
    uint32_t cap;
    int groups, bshift, cw, ch;
    size_t lds, o_window, o_hist, o_partials, total;
};

static int riw_plan(const dsx_points_desc *d, const dsx_view *v, RiwPlan &P, bool packed = false,
                    bool thr = false) {
    const int ncat = d->cat.kind == DSX_CAT_NONE ? 1 : d->cat.ncat;
    // bytes a window cell (packed with a threshold: the previous slot's key beside it)
    const size_t cb = packed ? (thr ? 16 : 8) : 4;
    P.lds = (size_t)kLdsBytes - 1024;  // the window, less the kernel's static words
    P.cap = (uint32_t)(P.lds / cb);
    if ((int64_t)P.cap / ncat < 1) return DSX_ERR_SHAPE;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && c > 0)
            cus = c;
    }
    const int64_t chunks = (d->n + kRiwChunk - 1) / kRiwChunk;
    P.groups = (int)(chunks < cus ? (chunks < 1 ? 1 : chunks) : cus);
    P.bshift = 0;
    while ((int64_t)((v->width + (1 << P.bshift) - 1) >> P.bshift) *
               ((v->height + (1 << P.bshift) - 1) >> P.bshift) > kSelBins)
        ++P.bshift;
    P.cw = (v->width + (1 << P.bshift) - 1) >> P.bshift;
    P.ch = (v->height + (1 << P.bshift) - 1) >> P.bshift;
    size_t o = 0;
    P.o_window = o; o = align_up(o + 4 * kRecWords, 256);
    P.o_hist = o; o = align_up(o + 4 * (size_t)kSelBins, 256);
    P.o_partials = o; o = align_up(o + cb * (size_t)P.cap * (size_t)P.groups, 256);
    P.total = o;
    return DSX_OK;
}

// NANCHK: rows whose value is NaN are skipped (the value column is read); CK, LIN: as
// in scan_kernel (the category read with the row's other columns, the branch-free linear
// projection); PK: the packed ARGMAX / ARGMIN form (float32 values).
template <typename TC, typename TV, bool NANCHK, int CK, bool LIN, bool PK = false, bool WIDE = false>
__global__ __launch_bounds__(kRiwThreads) void rowindex_window_kernel(RiwArgs r) {
    static_assert(kRiwThreads == kScanThreads && kRiwItems == kItems, "chunks as the scan's");
    static_assert(!PK || sizeof(TV) == 4, "packed: float32 values");
    extern __shared__ __align__(16) uint32_t win[];
    unsigned long long *const win64 = reinterpret_cast<unsigned long long *>(win);
    unsigned long long *const thr64 = win64 + r.cap;  // packed with a threshold
    const PointArgs &a = r.a;
    const int ncat = a.cat.ncat;
    const int tid = threadIdx.x;
    const Window w = load_window(r.window, a.W, a.H, (int64_t)(r.cap / (uint32_t)ncat));
    const uint32_t used = (uint32_t)(w.w * w.h * ncat);
    const uint32_t empty = r.last ? 0u : 0xFFFFFFFFu;
    const unsigned long long empty64 = r.last ? 0ull : ~0ull;
    const bool thr = PK && r.thr && r.kslot > 0;
    // packed: a is kept over b (the larger key wins when last, else the smaller)
    auto better = [&](unsigned long long x, unsigned long long y) { return r.last ? x > y : x < y; };
    for (uint32_t c = tid; c < r.cap; c += kRiwThreads) {
        if constexpr (PK) {
            win64[c] = empty64;
            if (thr && c < used) {  // the window cell's previous slot
                const uint32_t pix = c / (uint32_t)ncat, cat = c - pix * (uint32_t)ncat;
                const int32_t ly = (int32_t)(pix / (uint32_t)w.w), lx = (int32_t)(pix - (uint32_t)ly * w.w);
                const int64_t g = ((int64_t)(w.y0 + ly) * a.W + (w.x0 + lx)) * ncat + cat;
                thr64[c] = r.index[g * r.nsel + r.kslot - 1];
            }
        } else {
            win[c] = empty;
        }
    }
    __syncthreads();
    const int64_t nchunks = (a.n + kChunk - 1) / kChunk;
    // FIRST visits the chunks in increasing row order and LAST in decreasing order, so
    // that after the first sweep most rows fail the global skip test; a step past the
    // last chunk reads chunk nchunks (rows past the end: re-reads the last row, dropped)
    auto chunk_of = [&](int64_t k) { return k >= nchunks ? nchunks : (r.last ? nchunks - 1 - k : k); };
    using R = Rows<TC, TV, NANCHK, CK, WIDE>;
    auto *const index = global_ptr(r.index);
    auto *const sindex = global_ptr(reinterpret_cast<long long *>(r.index));
    // Three register sets rotate roles (binned / in flight / being loaded), as in the
    // points scan.  A row outside the window reads its cell of the index plane (a relaxed
    // device-scope load, served by L2) before any atomic: most rows cannot improve the
    // cell.  Those reads are issued before the loads of the chunk two ahead and waited for
    // after them, so that wait (vector memory counts complete in order) does not also wait
    // for those loads.
    auto step = [&](int64_t k, const R &S, R &far) {
        const int64_t ck = chunk_of(k);
        const int64_t b = ck * kChunk;
        const int64_t rem64 = a.n - b;
        const uint32_t rem = rem64 >= kChunk ? (uint32_t)kChunk : (rem64 > 0 ? (uint32_t)rem64 : 0u);
        int64_t cell[kRiwItems];
        long long cur[kRiwItems];
        unsigned long long pkv[kRiwItems];  // PK: the row's packed value
        unsigned long long pth[kRiwItems];  // PK with a threshold: the previous slot's key
        bool out[kRiwItems];
#pragma unroll
        for (int it = 0; it < kRiwItems; ++it) {
            // the row's place in the chunk: as the loader placed it
            const uint32_t j = WIDE ? (uint32_t)(4 * tid + it) : (uint32_t)(it * kRiwThreads + tid);
            int32_t px = 0, py = 0, pc = 0;
            bool keep;
            if constexpr (LIN) {  // as in scan_kernel
                const double x = (double)S.x[it], y = (double)S.y[it];
                const bool inr = (a.xmin <= x) & (x <= a.xmax) & (a.ymin <= y) & (y <= a.ymax);
                const double fx = inr ? __dadd_rn(__dmul_rn(x, a.sx), a.tx) : 0.0;
                const double fy = inr ? __dadd_rn(__dmul_rn(y, a.sy), a.ty) : 0.0;
                const int32_t xx = (int32_t)fx, yy = (int32_t)fy;
                px = xx < a.W - 1 ? xx : a.W - 1;
                py = yy < a.H - 1 ? yy : a.H - 1;
                keep = (j < rem) & inr;
            } else {
                keep = j < rem && locate_xy<TC>(a, S.x[it], S.y[it], px, py);
            }
            if constexpr (CK != CK_NONE) {
                pc = scan_category<CK>(a.cat, S.c[it]);
                keep = keep && pc >= 0;
            }
            if constexpr (NANCHK) keep = keep && !(S.v[it] != S.v[it]);
            out[it] = false;
            cell[it] = 0;
            cur[it] = 0;
            const uint32_t row = (uint32_t)(b + j);  // rows < 2^32 (checked on the host)
            pkv[it] = 0ull;
            pth[it] = 0ull;
            if constexpr (PK)
                pkv[it] = r.pkmode == 0 ? ((unsigned long long)fkey32((float)S.v[it]) << 32) | (r.last ? ~row : row)
                                        : (unsigned long long)(r.pkmode == 1 ? row : ~row);
            if (!keep) continue;
            if (in_window(w, px, py)) {
                const uint32_t lc = (uint32_t)(((py - w.y0) * w.w + (px - w.x0)) * ncat + pc);
                if constexpr (PK) {
                    if (thr && !better(thr64[lc], pkv[it])) continue;  // an earlier slot's
                    if (r.last) atomicMax(&win64[lc], pkv[it]);
                    else atomicMin(&win64[lc], pkv[it]);
                } else {
                    if (r.last) atomicMax(&win[lc], row + 1u);
                    else atomicMin(&win[lc], row);
                }
                continue;
            }
            out[it] = true;
            cell[it] = (((int64_t)py * a.W + px) * ncat + pc) * r.nsel + r.kslot;  // its slot
        }
        // (global-memory pointers: through a generic one these would be flat operations,
        // whose waits also wait for LDS)
#pragma unroll
        for (int it = 0; it < kRiwItems; ++it) {
            if (out[it]) cur[it] = (long long)__hip_atomic_load(index + cell[it], __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_AGENT);
            if (thr && out[it]) pth[it] = __hip_atomic_load(index + cell[it] - 1, __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_AGENT);
        }
        far.load(a, chunk_of(k + 2 * (int64_t)gridDim.x), tid);
#pragma unroll
        for (int it = 0; it < kRiwItems; ++it) {
            if (!out[it]) continue;
            if constexpr (PK) {  // unsigned: the plane starts at 0 (ARGMAX) / all ones (ARGMIN)
                const unsigned long long c_ = (unsigned long long)cur[it];
                if (thr && !better(pth[it], pkv[it])) continue;  // an earlier slot's
                if (r.last) {
                    if (pkv[it] > c_)
                        __hip_atomic_fetch_max(index + cell[it], pkv[it], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                } else if (pkv[it] < c_) {
                    __hip_atomic_fetch_min(index + cell[it], pkv[it], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                }
                continue;
            }
            const long long row =
                (long long)(b + (WIDE ? 4 * tid + it : it * kRiwThreads + tid));
            if (r.last) {
                if (row > cur[it])
                    __hip_atomic_fetch_max(sindex + cell[it], row, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            } else if ((unsigned long long)row < (unsigned long long)cur[it]) {
                // -1 (empty) is the largest unsigned value
                __hip_atomic_fetch_min(index + cell[it], (unsigned long long)row, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    };
    R sa, sb, sc;
    const int64_t G = gridDim.x;
    sa.load(a, chunk_of(blockIdx.x), tid);
    sb.load(a, chunk_of(blockIdx.x + G), tid);
    for (int64_t k = blockIdx.x; k < nchunks; k += 3 * G) {  // steps past the end drop rows
        step(k, sa, sc);
        step(k + G, sb, sa);
        step(k + 2 * G, sc, sb);
    }
    __syncthreads();
    if constexpr (PK) {
        unsigned long long *mine = reinterpret_cast<unsigned long long *>(r.partials) +
                                   (size_t)blockIdx.x * r.cap;
        for (uint32_t c = tid; c < used; c += kRiwThreads) mine[c] = win64[c];
    } else {
        uint32_t *mine = r.partials + (size_t)blockIdx.x * r.cap;
        for (uint32_t c = tid; c < used; c += kRiwThreads) mine[c] = win[c];
    }
}

// The window cells of the index plane: min (FIRST) / max (LAST) over the workgroups.
__global__ __launch_bounds__(256) void rowindex_window_merge(RiwArgs r, int groups) {
    const PointArgs &a = r.a;
    const int ncat = a.cat.ncat;
    const Window w = load_window(r.window, a.W, a.H, (int64_t)(r.cap / (uint32_t)ncat));
    const uint32_t used = (uint32_t)(w.w * w.h * ncat);
    if (r.packed) {  // the packed plane's window cells: max (ARGMAX) / min (ARGMIN)
        const unsigned long long *part = reinterpret_cast<const unsigned long long *>(r.partials);
        for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < used; c += gridDim.x * blockDim.x) {
            unsigned long long m = r.last ? 0ull : ~0ull;
            int q = 0;
            for (; q + 8 <= groups; q += 8) {
                unsigned long long x[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) x[u] = part[(size_t)(q + u) * r.cap + c];
#pragma unroll
                for (int u = 0; u < 8; ++u) m = r.last ? (x[u] > m ? x[u] : m) : (x[u] < m ? x[u] : m);
            }
            for (; q < groups; ++q) {
                const unsigned long long x = part[(size_t)q * r.cap + c];
                m = r.last ? (x > m ? x : m) : (x < m ? x : m);
            }
            const uint32_t pix = c / (uint32_t)ncat, cat = c - pix * (uint32_t)ncat;
            const int32_t ly = (int32_t)(pix / (uint32_t)w.w), lx = (int32_t)(pix - (uint32_t)ly * w.w);
            r.index[(((int64_t)(w.y0 + ly) * a.W + (w.x0 + lx)) * ncat + cat) * r.nsel + r.kslot] = m;
        }
        return;
    }
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < used; c += gridDim.x * blockDim.x) {
        uint32_t m = r.last ? 0u : 0xFFFFFFFFu;
        auto fold = [&](uint32_t x) { m = r.last ? (x > m ? x : m) : (x < m ? x : m); };
        int q = 0;
        for (; q + 8 <= groups; q += 8) {  // eight groups' loads in flight at a time
            uint32_t x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = r.partials[(size_t)(q + u) * r.cap + c];
#pragma unroll
            for (int u = 0; u < 8; ++u) fold(x[u]);
        }
        for (; q < groups; ++q) fold(r.partials[(size_t)q * r.cap + c]);
        const uint32_t pix = c / (uint32_t)ncat, cat = c - pix * (uint32_t)ncat;
        const int32_t ly = (int32_t)(pix / (uint32_t)w.w), lx = (int32_t)(pix - (uint32_t)ly * w.w);
        const int64_t g = ((int64_t)(w.y0 + ly) * a.W + (w.x0 + lx)) * ncat + cat;
        long long out;
        if (r.last) out = m == 0u ? -1ll : (long long)(m - 1u);
        else out = m == 0xFFFFFFFFu ? -1ll : (long long)m;
        r.index[g] = (unsigned long long)out;
    }
}

// The packed plane to rows: the low word (ARGMIN) or its complement (ARGMAX), -1 where no
// row was kept (the plane's start value).
__global__ __launch_bounds__(256) void rowindex_unpack_kernel(unsigned long long *p, int64_t n, int amax) {
    const unsigned long long empty = amax ? 0ull : ~0ull;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long v = p[i];
        const uint32_t lo32 = (uint32_t)v;
        p[i] = v == empty ? ~0ull : (unsigned long long)(amax ? ~lo32 : lo32);
    }
}

// pick: set the call's window record first (else it is already in the workspace);
// unpack: turn a packed plane into rows afterwards.
template <typename TC>
static int riw_launch(const dsx_points_desc *d, const RiwArgs &r, const RiwPlan &P, uint8_t *ws,
                      hipStream_t st, bool pick = true, bool unpack = true) {
    int32_t *window = const_cast<int32_t *>(r.window);
    uint32_t *hist_sel = reinterpret_cast<uint32_t *>(ws + P.o_hist);
    int s;
    if (!pick) {
    } else if (d->window && d->reuse_window) {
        // the caller's stored window (sanitised by every kernel that reads it)
        hipLaunchKernelGGL(copy_window_kernel, dim3(1), dim3(64), 0, st,
                           (const int32_t *)d->window, window, (int32_t)DSX_STRATEGY_WINDOW);
    } else {
        if ((s = fill_async(hist_sel, DSX_FILL_ZERO32, kSelBins, st))) return s;
        hipLaunchKernelGGL((select_hist_kernel<TC>), dim3(64), dim3(1024), 0, st, r.a, P.bshift,
                           P.cw, P.ch, hist_sel);
        hipLaunchKernelGGL(select_pick_kernel, dim3(1), dim3(1024), 0, st,
                           (const uint32_t *)hist_sel, P.bshift, P.cw, P.ch, r.a.W, r.a.H,
                           (int64_t)(P.cap / (uint32_t)r.a.cat.ncat), 0, window);
        if (d->window)
            hipLaunchKernelGGL(copy_window_kernel, dim3(1), dim3(64), 0, st,
                               (const int32_t *)window, d->window, 0);
    }
    if ((s = launch_status())) return s;
    // (the linear fast path needs the reference's own translation, as in the points scan)
    const PointArgs &a = r.a;
    const bool lin = !a.xlog && !a.ylog && a.sx > 0 && a.sy > 0 && a.tx == -(a.xmin * a.sx) &&
                     a.ty == -(a.ymin * a.sy);
    const bool f32v = d->val_dtype == DSX_F32;
    const bool nan = r.nan_check != 0;
    const bool byte = (a.cat.dtype == DSX_I8 || a.cat.dtype == DSX_U8) && a.cat.kind == DSX_CAT_CODES;
    auto kern = rowindex_window_kernel<TC, double, false, CK_ANY, false>;
    if (a.cat.kind == DSX_CAT_NONE) {
        if (lin) kern = nan ? (f32v ? rowindex_window_kernel<TC, float, true, CK_NONE, true>
                                    : rowindex_window_kernel<TC, double, true, CK_NONE, true>)
                            : rowindex_window_kernel<TC, double, false, CK_NONE, true>;
        else kern = nan ? (f32v ? rowindex_window_kernel<TC, float, true, CK_NONE, false>
                                : rowindex_window_kernel<TC, double, true, CK_NONE, false>)
                        : rowindex_window_kernel<TC, double, false, CK_NONE, false>;
    } else if (byte) {
        kern = nan ? (f32v ? rowindex_window_kernel<TC, float, true, CK_CODE8, false>
                           : rowindex_window_kernel<TC, double, true, CK_CODE8, false>)
                   : rowindex_window_kernel<TC, double, false, CK_CODE8, false>;
    } else {
        kern = nan ? (f32v ? rowindex_window_kernel<TC, float, true, CK_ANY, false>
                           : rowindex_window_kernel<TC, double, true, CK_ANY, false>)
                   : rowindex_window_kernel<TC, double, false, CK_ANY, false>;
    }
    if (r.packed && nan) {  // float32 values (checked by the caller), NaN rows skipped
        if (a.cat.kind == DSX_CAT_NONE)
            kern = lin ? rowindex_window_kernel<TC, float, true, CK_NONE, true, true>
                       : rowindex_window_kernel<TC, float, true, CK_NONE, false, true>;
        else if (byte) kern = rowindex_window_kernel<TC, float, true, CK_CODE8, false, true>;
        else kern = rowindex_window_kernel<TC, float, true, CK_ANY, false, true>;
    } else if (r.packed) {  // row keys, no value read
        if (a.cat.kind == DSX_CAT_NONE)
            kern = lin ? rowindex_window_kernel<TC, float, false, CK_NONE, true, true>
                       : rowindex_window_kernel<TC, float, false, CK_NONE, false, true>;
        else if (byte) kern = rowindex_window_kernel<TC, float, false, CK_CODE8, false, true>;
        else kern = rowindex_window_kernel<TC, float, false, CK_ANY, false, true>;
    }
    // float32 columns, 16-byte aligned, whole groups of four rows, linear axes, no
    // categorizer: the wide row loads
    if constexpr (sizeof(TC) == 4) {
        const bool wide = lin && a.cat.kind == DSX_CAT_NONE && d->n >= 4 && d->n % 4 == 0 &&
                          ((uintptr_t)a.x % 16) == 0 && ((uintptr_t)a.y % 16) == 0 &&
                          (!nan || (f32v && ((uintptr_t)a.val % 16) == 0));
        if (wide) {
            if (r.packed)
                kern = nan ? rowindex_window_kernel<TC, float, true, CK_NONE, true, true, true>
                           : rowindex_window_kernel<TC, float, false, CK_NONE, true, true, true>;
            else
                kern = nan ? rowindex_window_kernel<TC, float, true, CK_NONE, true, false, true>
                           : rowindex_window_kernel<TC, double, false, CK_NONE, true, false, true>;
        }
    }
    const hipError_t e = hipFuncSetAttribute((const void *)kern,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)P.lds);
    if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
    hipLaunchKernelGGL(kern, dim3(P.groups), dim3(kRiwThreads), P.lds, st, r);
    if ((s = launch_status())) return s;
    hipLaunchKernelGGL(rowindex_window_merge, dim3(grid_for(P.cap, 256, 4)), dim3(256), 0, st, r,
                       P.groups);
    if (r.packed && unpack) {
        if ((s = launch_status())) return s;
        const int64_t ncells = (int64_t)a.W * a.H * a.cat.ncat;
        hipLaunchKernelGGL(rowindex_unpack_kernel, dim3(grid_for(ncells, 256, 4)), dim3(256), 0, st,
                           r.index, ncells, r.last);
    }
    return launch_status();
}

// ARGMAX / ARGMIN take the packed window pass when no key plane is given, with a
// workspace, over a float32 value column.
static bool ri_packed(const dsx_points_desc *d, const dsx_rowindex_desc *r) {
    return (r->mode == DSX_RI_ARGMAX || r->mode == DSX_RI_ARGMIN) && !r->key && has_value(d) &&
           d->val_dtype == DSX_F32;
}

extern "C" size_t dsx_points_rowindex_workspace_size(const dsx_points_desc *d, const dsx_view *v,
                                                     const dsx_rowindex_desc *r) {
    if (check_points(d, v) || !r) return 0;
    RiwPlan P;
    if (riw_plan(d, v, P, ri_packed(d, r))) return 0;
    return P.total;
}

extern "C" int dsx_points_rowindex(const dsx_points_desc *d, const dsx_view *v,
                                   const dsx_rowindex_desc *r, void *workspace,
                                   size_t workspace_bytes, void *stream) {
    int st = check_points(d, v);
    if (st) return st;
    if (!r || !r->index) return DSX_ERR_NULL;
    if (r->mode < DSX_RI_FIRST || r->mode > DSX_RI_ARGMIN) return DSX_ERR_MODE;
    const bool arg = r->mode == DSX_RI_ARGMAX || r->mode == DSX_RI_ARGMIN;
    if ((r->nan_check || arg) && !has_value(d)) return DSX_ERR_ACCUM;
    const bool packed = ri_packed(d, r) && workspace;
    if (arg && !r->key && !packed) return DSX_ERR_NULL;
    const int64_t ncells = (int64_t)v->width * v->height * cat_of(d->cat).ncat;
    hipStream_t s = as_stream(stream);
    // (the packed ARGMAX plane starts at 0, below every packed value)
    const int fill = packed && r->mode == DSX_RI_ARGMAX ? DSX_FILL_ZERO64 : DSX_FILL_ONES64;
    if ((st = fill_async(r->index, fill, ncells, s))) return st;
    if (d->n == 0 && packed) return fill_async(r->index, DSX_FILL_ONES64, ncells, s);
    if (d->n == 0) return DSX_OK;
    RiArgs ra;
    ra.a = args_of(d, v);
    ra.mode = r->mode;
    ra.nan_check = (r->nan_check || arg) ? 1 : 0;
    ra.key = r->key;
    ra.index = reinterpret_cast<unsigned long long *>(r->index);
    const int vdt = has_value(d) ? d->val_dtype : DSX_F64;
    if ((!arg && workspace) || packed) {
        // FIRST / LAST with a workspace, or packed ARGMAX / ARGMIN: the LDS-window pass
        RiwPlan P;
        if ((st = riw_plan(d, v, P, packed))) return st;
        if (workspace_bytes < P.total) return DSX_ERR_WORKSPACE;
        uint8_t *ws = reinterpret_cast<uint8_t *>(workspace);
        RiwArgs rw;
        rw.a = ra.a;
        rw.last = (r->mode == DSX_RI_LAST || r->mode == DSX_RI_ARGMAX) ? 1 : 0;
        rw.packed = packed ? 1 : 0;
        rw.pkmode = 0;
        rw.nsel = 1;
        rw.kslot = 0;
        rw.thr = 0;
        rw.nan_check = ra.nan_check;
        rw.index = ra.index;
        rw.window = reinterpret_cast<const int32_t *>(ws + P.o_window);
        rw.cap = P.cap;
        rw.partials = reinterpret_cast<uint32_t *>(ws + P.o_partials);
        if (!has_value(d)) { rw.a.val = rw.a.x; rw.nan_check = 0; }
        dsx_points_desc dd = *d;
        if (!has_value(d)) dd.val_dtype = d->xy_dtype;
        if (d->xy_dtype == DSX_F32) return riw_launch<float>(&dd, rw, P, ws, s);
        return riw_launch<double>(&dd, rw, P, ws, s);
    }
    if (d->xy_dtype == DSX_F32) return rowindex_launch<float>(ra, vdt, s);
    return rowindex_launch<double>(ra, vdt, s);
}

// The n-pass form of the n-value reductions (called by dsx_points_nselect): pass k keeps
// slot k of every cell with the packed window pass, the previous slot's key the bound a
// row's key must pass.  Packed keys: MAX key(v) << 32 | ~row (largest wins), MIN
// key(v) << 32 | row, FIRST row, LAST ~row (smallest wins).  The slots keep their packed
// keys (the caller turns them into rows).  Float32 values for MAX / MIN and for a NaN
// check; the caller fills index with the empty key (0 for MAX, else all ones).
namespace dsx {
size_t ri_npass_workspace_size(const dsx_points_desc *d, const dsx_view *v) {
    RiwPlan P;
    if (riw_plan(d, v, P, true, true)) return 0;
    return P.total;
}

int ri_npass(const dsx_points_desc *d, const dsx_view *v, int mode, int nsel, int nan_check,
             int64_t *index, void *workspace, size_t workspace_bytes, hipStream_t st) {
    RiwPlan P;
    int s;
    if ((s = riw_plan(d, v, P, true, true))) return s;
    if (!workspace || workspace_bytes < P.total) return DSX_ERR_WORKSPACE;
    uint8_t *ws = reinterpret_cast<uint8_t *>(workspace);
    const bool value = mode == DSX_NS_MAX || mode == DSX_NS_MIN;
    RiwArgs rw;
    rw.a = args_of(d, v);
    rw.last = mode == DSX_NS_MAX ? 1 : 0;
    rw.nan_check = (value || nan_check) ? 1 : 0;
    rw.index = reinterpret_cast<unsigned long long *>(index);
    rw.window = reinterpret_cast<const int32_t *>(ws + P.o_window);
    rw.cap = P.cap;
    rw.partials = reinterpret_cast<uint32_t *>(ws + P.o_partials);
    rw.packed = 1;
    rw.pkmode = value ? 0 : (mode == DSX_NS_FIRST ? 1 : 2);
    rw.nsel = nsel;
    rw.thr = 1;
    if (!rw.nan_check) rw.a.val = rw.a.x;  // (no value read)
    dsx_points_desc dd = *d;
    dd.window = nullptr;  // the pass picks its own window, once
    dd.reuse_window = 0;
    for (int k = 0; k < nsel; ++k) {
        rw.kslot = k;
        s = d->xy_dtype == DSX_F32 ? riw_launch<float>(&dd, rw, P, ws, st, k == 0, false)
                                   : riw_launch<double>(&dd, rw, P, ws, st, k == 0, false);
        if (s) return s;
    }
    return DSX_OK;
}
}  // namespace dsx

extern "C" size_t dsx_points_indexed_workspace_size(const dsx_points_desc *d, const dsx_view *v,
                                                    const dsx_planes *p, const dsx_grid_ref *g) {
    if (idx_check(d, v, p, nullptr, g)) return 0;
    IdxPlan P;
    if (idx_make_plan(d, v, p, P)) return 0;
    return P.total;
}

extern "C" int dsx_points_agg_indexed(const dsx_points_desc *d, const dsx_view *v,
                                      const dsx_planes *p, const double *mean,
                                      const dsx_grid_ref *g, void *workspace,
                                      size_t workspace_bytes, void *stream) {
    int st = idx_check(d, v, p, mean, g);
    if (st) return st;
    IdxPlan P;
    if ((st = idx_make_plan(d, v, p, P))) return st;
    if (!workspace || workspace_bytes < P.total) return DSX_ERR_WORKSPACE;
    uint8_t *ws = reinterpret_cast<uint8_t *>(workspace);
    IdxArgs A;
    A.a = args_of(d, v);
    if (d->n == 0) {  // no rows: every tile reads nothing (the columns are never touched)
        A.a.x = A.a.y = A.a.val = g->cell_start;
    }
    A.L = P.L;
    A.tw = P.tw; A.th = P.th; A.ntx = P.ntx; A.nty = P.nty; A.ntiles = P.ntiles;
    A.grid = g->grid;
    A.ex0 = g->x0; A.esx = (double)g->grid / (g->x1 - g->x0);
    A.ey0 = g->y0; A.esy = (double)g->grid / (g->y1 - g->y0);
    A.cs = g->cell_start;
    A.tile_rows = reinterpret_cast<unsigned long long *>(ws + P.o_rows);
    A.part_start = reinterpret_cast<uint32_t *>(ws + P.o_pstart);
    A.part_tile = reinterpret_cast<uint32_t *>(ws + P.o_ptile);
    A.part_len = reinterpret_cast<unsigned long long *>(ws + P.o_plen);
    A.max_parts = P.max_parts;
    A.cus = P.cus;
    A.scratch = ws + P.o_scratch;
    A.mean = mean;
    A.out = planes_of(p);
    hipStream_t s = as_stream(stream);
    if (d->xy_dtype == DSX_F32) return idx_dispatch<float>(d, A, P, s);
    return idx_dispatch<double>(d, A, P, s);
}
