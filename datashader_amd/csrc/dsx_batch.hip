// dsx_batch.hip -- batched viewports over a resident, lattice-sorted point set: many
// Canvas.points requests (each its own ranges and canvas size) aggregated by one replay of
// a captured HIP graph.
//
// No counterpart in the reference, which scans every row per call (core.py:1334-1359 ->
// data_libraries/pandas.py:23-66); the per-row work is still glyphs/points.py:188-212
// (projection, range test, W-1 / H-1 clamp) with the count / count(v) / _sum_zero appends
// of reductions.py (554-558, 582-584, 958-962) and the sum / mean finalizers (1092-1096,
// 1294-1297), so each request gets the reference's result for the whole set.
//
// One batch is a fixed sequence, captured once per engine and replayed: a copy of the
// request table from pinned host memory into the workspace, then five kernels:
//   stage   expands the requests (tile grid of each), resets the work-queue counters and
//           maps every tile to its request
//   tiles   one workgroup per canvas tile: the lattice cells whose rows can land in the
//           tile's pixels (the tile's coordinate range, widened by a rounding margin,
//           mapped through the index's own cell formula -- monotone, so no cell is
//           missed), their row count, and the tile's parts (row ranges of at most
//           part_rows rows; at least 1)
//   items   one workgroup: exclusive scans of the parts -> the work queue (tile, part)
//           and the scratch regions of split tiles
//   agg     persistent workgroups pop work items from the queue (one atomic per item):
//           the tile's lattice rows are tabled in LDS, the part's rows are projected and
//           accumulated in an LDS copy of the tile's planes (LDS atomics), and the tile is
//           written with the finalizers applied (mean = sum / count(v), NaN where empty)
//           -- or, for a part of a split tile, left in scratch
//   merge   split tiles: the parts' regions folded, finalized and written
// Every output cell is written exactly once, so output planes need no initialisation.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <new>

#include "dsx_common.h"

namespace dsx {

constexpr int kBLds = 160 * 1024;
constexpr int kBMaxLatRows = 4096;            // lattice rows a tile may span: the largest grid
constexpr int kBTable = kBMaxLatRows * 8;     // LDS: first row + row prefix (u32) per lattice row
constexpr int kBTileThreads = 256;
constexpr int kBItemThreads = 1024;
constexpr int kBHeader = 256;                 // bytes of the table header

// planes an engine produces
enum : uint32_t { BP_COUNT = 1, BP_COUNTV = 2, BP_SUM = 4, BP_MEAN = 8 };

struct BHeader {           // at the start of the pinned host table
    int32_t nreq, ntiles;
};

// Region of one tile in LDS / scratch: sum f64 [cap], count u32 [cap], nan u32 [cap]
struct BLay {
    uint32_t cap, o_sum, o_cnt, o_nan, bytes;
    bool has_val;
};

struct BReq {              // a request as the kernels read it (device copy)
    double sx, tx, sy, ty, xmin, xmax, ymin, ymax;
    int32_t W, H, tw, th, ntx, nty, tile_base, pad;
    int64_t part_rows;
    uint32_t *count, *countv;
    double *sum, *mean;
};

struct BTile {
    int32_t req, t;
    int32_t gx0, gx1, gy0, gy1;    // lattice cells (inclusive); gx0 > gx1: no rows
    uint64_t rows;
    uint32_t part_base, nparts;    // work items [part_base, part_base + nparts)
    uint32_t region;               // first scratch region (split tiles)
    uint32_t pad;
};

struct BItem {
    uint32_t tile, part;
};

struct BArgs {
    const dsx_batch_req *host;     // the workspace's copy of the request table
    const BHeader *hhdr;           // ... and of its header
    BReq *reqs;
    BTile *tiles;
    int32_t *tile_req;
    BItem *items;
    uint32_t *split;               // split tile ids
    uint32_t *ctr;                 // [0] queue head, [1] items, [2] split tiles, [3] regions,
                                   // [4] requests, [5] tiles (the header's, staged)
    uint8_t *scratch;
    int32_t max_req, max_tiles, max_items, max_regions;
    BLay L;
    int64_t n;
    const float *x, *y, *v;
    int32_t grid;
    double ex0, esx, ey0, esy;
    const int64_t *cs;
    uint32_t planes;
};

__device__ __forceinline__ int32_t b_cell(double v, double e0, double es, int32_t grid) {
    // the index's own formula (dsx_serve.hip grid_cell): monotone non-decreasing in v
    const double top = (double)(grid - 1);
    double f = (v - e0) * es;
    f = f < 0.0 ? 0.0 : (f > top ? top : f);
    return (int32_t)f;
}

// Coordinate range [lo, hi] holding every value that projects into pixels [p0, p1) of an
// axis (scale s > 0, translation t, data bounds [vmin, vmax], n pixels): int(v*s + t) in
// [p0, p1), or >= n for the last pixel.  Widened by a relative margin far above the
// rounding of v*s + t, so that the cells of [lo, hi] hold every such row.
__device__ __forceinline__ void b_range(int32_t p0, int32_t p1, int32_t n, double s, double t,
                                        double vmin, double vmax, double &lo, double &hi) {
    lo = ((double)p0 - t) / s;
    hi = ((double)p1 - t) / s;
    const double m = 1e-12 * (fabs(lo) + fabs(hi) + fabs(vmin) + fabs(vmax) + 1.0 / s);
    lo -= m;
    hi += m;
    lo = lo > vmin ? lo : vmin;
    hi = (p1 >= n || hi > vmax) ? vmax : hi;
}

__device__ __forceinline__ void b_tile_rect(const BReq &r, int32_t t, int32_t &px0, int32_t &px1,
                                            int32_t &py0, int32_t &py1) {
    const int32_t tx = t % r.ntx, ty = t / r.ntx;
    px0 = tx * r.tw;
    px1 = px0 + r.tw < r.W ? px0 + r.tw : r.W;
    py0 = ty * r.th;
    py1 = py0 + r.th < r.H ? py0 + r.th : r.H;
}

template <typename T>
__device__ __forceinline__ T b_block_exscan(T v, T *sh, T &total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    T x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const T y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            const T c = sh[k];
            sh[k] = run;
            run += c;
        }
        sh[blockDim.x >> 6] = run;
    }
    __syncthreads();
    total = sh[blockDim.x >> 6];
    const T r = sh[wv] + x - v;
    __syncthreads();
    return r;
}

// ---- stage: table -> workspace, counters reset, tile -> request map --------------------
__global__ __launch_bounds__(256) void b_stage_kernel(BArgs A) {
    __shared__ dsx_batch_req sq;
    const int32_t nreq = A.hhdr->nreq;
    if (blockIdx.x == 0 && threadIdx.x < 4) A.ctr[threadIdx.x] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 4) A.ctr[4] = (uint32_t)nreq;
    if (blockIdx.x == 0 && threadIdx.x == 5) A.ctr[5] = (uint32_t)A.hhdr->ntiles;
    for (int32_t r = blockIdx.x; r < nreq && r < A.max_req; r += gridDim.x) {
        if (threadIdx.x == 0) sq = A.host[r];
        __syncthreads();
        const dsx_batch_req q = sq;
        __syncthreads();
        const int32_t ntx = (q.width + q.tile_w - 1) / q.tile_w;
        const int32_t nty = (q.height + q.tile_h - 1) / q.tile_h;
        if (threadIdx.x == 0) {
            BReq d;
            d.sx = q.sx; d.tx = q.tx; d.sy = q.sy; d.ty = q.ty;
            d.xmin = q.xmin; d.xmax = q.xmax; d.ymin = q.ymin; d.ymax = q.ymax;
            d.W = q.width; d.H = q.height; d.tw = q.tile_w; d.th = q.tile_h;
            d.ntx = ntx; d.nty = nty; d.tile_base = q.tile_base; d.pad = 0;
            d.part_rows = q.part_rows;
            d.count = q.count; d.countv = q.countv; d.sum = q.sum; d.mean = q.mean;
            A.reqs[r] = d;
        }
        const int32_t nt = ntx * nty;
        for (int32_t t = threadIdx.x; t < nt; t += blockDim.x)
            if (q.tile_base + t < A.max_tiles) A.tile_req[q.tile_base + t] = r;
    }
}

// ---- tiles: lattice cells, rows and parts of every tile --------------------------------
__global__ __launch_bounds__(kBTileThreads) void b_tiles_kernel(BArgs A) {
    __shared__ unsigned long long red[kBTileThreads / 64 + 1];
    const int32_t ntiles = (int32_t)A.ctr[5];
    for (int32_t b = blockIdx.x; b < ntiles && b < A.max_tiles; b += gridDim.x) {
        const int32_t r = A.tile_req[b];
        const BReq q = A.reqs[r];
        const int32_t t = b - q.tile_base;
        int32_t px0, px1, py0, py1;
        b_tile_rect(q, t, px0, px1, py0, py1);
        double xl, xh, yl, yh;
        b_range(px0, px1, q.W, q.sx, q.tx, q.xmin, q.xmax, xl, xh);
        b_range(py0, py1, q.H, q.sy, q.ty, q.ymin, q.ymax, yl, yh);
        int32_t gx0 = 1, gx1 = 0, gy0 = 1, gy1 = 0;  // empty unless the range is non-empty
        if (xl <= xh && yl <= yh) {
            gx0 = b_cell(xl, A.ex0, A.esx, A.grid);
            gx1 = b_cell(xh, A.ex0, A.esx, A.grid);
            gy0 = b_cell(yl, A.ey0, A.esy, A.grid);
            gy1 = b_cell(yh, A.ey0, A.esy, A.grid);
        }
        unsigned long long sum = 0;
        if (gx0 <= gx1)
            for (int32_t gy = gy0 + (int32_t)threadIdx.x; gy <= gy1; gy += blockDim.x) {
                const int64_t row = (int64_t)gy * A.grid;
                sum += (unsigned long long)(A.cs[row + gx1 + 1] - A.cs[row + gx0]);
            }
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long s2 = 0;
            for (int k = 0; k < kBTileThreads / 64; ++k) s2 += red[k];
            BTile d;
            d.req = r; d.t = t;
            d.gx0 = gx0; d.gx1 = gx1; d.gy0 = gy0; d.gy1 = gy1;
            d.rows = s2;
            const unsigned long long E = (unsigned long long)(q.part_rows > 0 ? q.part_rows : 1);
            unsigned long long np = (s2 + E - 1) / E;
            d.nparts = (uint32_t)(np < 1 ? 1 : np);
            d.part_base = 0; d.region = 0; d.pad = 0;
            A.tiles[b] = d;
        }
        __syncthreads();
    }
}

// ---- items: the work queue ------------------------------------------------------------
// Parts per tile as planned, unless the split parts would exceed the scratch regions or
// the items the queue holds: then every tile's part length grows to
// max(part_rows, 2 * rows / regions), which bounds the split parts by the regions
// (a tile of k > 1 parts has rows > (k - 1) * E, so sum(k) over split tiles < 2 rows / E).
__global__ __launch_bounds__(kBItemThreads) void b_items_kernel(BArgs A) {
    __shared__ unsigned long long sh[kBItemThreads / 64 + 1];
    __shared__ unsigned long long s_E;
    const int32_t ntiles = (int32_t)A.ctr[5] < A.max_tiles ? (int32_t)A.ctr[5] : A.max_tiles;
    unsigned long long rows = 0, split = 0, items = 0;
    for (int32_t b = threadIdx.x; b < ntiles; b += blockDim.x) {
        const BTile &d = A.tiles[b];
        rows += d.rows;
        items += d.nparts;
        if (d.nparts > 1) split += d.nparts;
    }
    unsigned long long tot;
    b_block_exscan(rows, sh, tot);
    const unsigned long long total_rows = tot;
    b_block_exscan(split, sh, tot);
    const unsigned long long total_split = tot;
    b_block_exscan(items, sh, tot);
    const unsigned long long total_items = tot;
    if (threadIdx.x == 0) {
        unsigned long long E = 0;  // 0: keep every tile's planned parts
        if (total_split > (unsigned long long)A.max_regions ||
            total_items > (unsigned long long)A.max_items) {
            const unsigned long long regions = (unsigned long long)A.max_regions;
            const unsigned long long items_room =
                (unsigned long long)(A.max_items > ntiles ? A.max_items - ntiles : 1);
            const unsigned long long room = regions < items_room ? regions : items_room;
            E = (2 * total_rows + room - 1) / (room > 0 ? room : 1);
            E = E < 1 ? 1 : E;
        }
        s_E = E;
    }
    __syncthreads();
    const unsigned long long Eall = s_E;
    unsigned long long icarry = 0, rcarry = 0, scarry = 0;
    for (int32_t b0 = 0; b0 < ntiles; b0 += blockDim.x) {
        const int32_t b = b0 + (int32_t)threadIdx.x;
        unsigned long long np = 0;
        if (b < ntiles) {
            BTile &d = A.tiles[b];
            if (Eall) {
                const unsigned long long own = (unsigned long long)A.reqs[d.req].part_rows;
                const unsigned long long E = own > Eall ? own : Eall;
                unsigned long long k = (d.rows + E - 1) / E;
                d.nparts = (uint32_t)(k < 1 ? 1 : k);
            }
            np = d.nparts;
        }
        unsigned long long tot_i, tot_r, tot_s;
        const unsigned long long ex_i = b_block_exscan(np, sh, tot_i);
        const unsigned long long ex_r = b_block_exscan(np > 1 ? np : 0ull, sh, tot_r);
        const unsigned long long ex_s = b_block_exscan(np > 1 ? 1ull : 0ull, sh, tot_s);
        if (b < ntiles) {
            BTile &d = A.tiles[b];
            d.part_base = (uint32_t)(icarry + ex_i);
            d.region = (uint32_t)(rcarry + ex_r);
            for (unsigned long long q = 0; q < np; ++q) {
                const unsigned long long it = icarry + ex_i + q;
                if (it < (unsigned long long)A.max_items) {
                    BItem m;
                    m.tile = (uint32_t)b;
                    m.part = (uint32_t)q;
                    A.items[it] = m;
                }
            }
            if (np > 1 && scarry + ex_s < (unsigned long long)A.max_tiles)
                A.split[scarry + ex_s] = (uint32_t)b;
        }
        icarry += tot_i;
        rcarry += tot_r;
        scarry += tot_s;
    }
    if (threadIdx.x == 0) {
        A.ctr[1] = (uint32_t)(icarry < (unsigned long long)A.max_items ? icarry : A.max_items);
        A.ctr[2] = (uint32_t)scarry;
        A.ctr[3] = (uint32_t)rcarry;
    }
}

// ---- agg: persistent workgroups over the work queue ---------------------------------
__device__ __forceinline__ void b_region_init(uint8_t *r, const BLay &L, int tid, int nt) {
    for (uint32_t c = tid; c < L.cap; c += nt) {
        reinterpret_cast<uint32_t *>(r + L.o_cnt)[c] = 0u;
        if (L.has_val) {
            reinterpret_cast<double *>(r + L.o_sum)[c] = 0.0;
            reinterpret_cast<uint32_t *>(r + L.o_nan)[c] = 0u;
        }
    }
}

// Write cell c of a finished region as the tile's output cell g: the finalizers of
// count(), count(v), sum(v) (NaN where no value) and mean(v).
__device__ __forceinline__ void b_store_cell(const uint8_t *r, const BLay &L, uint32_t c,
                                             const BReq &q, int64_t g) {
    const uint32_t n = reinterpret_cast<const uint32_t *>(r + L.o_cnt)[c];
    if (q.count) global_ptr(q.count)[g] = n;
    if (!L.has_val) return;
    const uint32_t nv = n - reinterpret_cast<const uint32_t *>(r + L.o_nan)[c];
    const double s = reinterpret_cast<const double *>(r + L.o_sum)[c];
    if (q.countv) global_ptr(q.countv)[g] = nv;
    if (q.sum) global_ptr(q.sum)[g] = nv > 0u ? s : (double)NAN;
    if (q.mean) global_ptr(q.mean)[g] = nv > 0u ? s / (double)nv : (double)NAN;
}

__device__ __forceinline__ void b_store_tile(const uint8_t *r, const BLay &L, const BReq &q,
                                             int32_t t, int nt) {
    int32_t px0, px1, py0, py1;
    b_tile_rect(q, t, px0, px1, py0, py1);
    const int32_t w = px1 - px0, h = py1 - py0;
    for (int32_t c = threadIdx.x; c < w * h; c += nt) {
        const int32_t ly = c / w, lx = c - ly * w;
        b_store_cell(r, L, (uint32_t)(ly * q.tw + lx), q, (int64_t)(py0 + ly) * q.W + (px0 + lx));
    }
}

constexpr int kBItems = 4;  // rows per thread per iteration, loads issued together

template <int NT>
__global__ __launch_bounds__(NT) void b_agg_kernel(BArgs A) {
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ unsigned long long sh[NT / 64 + 1];
    __shared__ uint32_t s_item;
    const BLay &L = A.L;
    // the tile's lattice rows: first sorted row and row-count prefix, 32-bit (n < 2^32)
    uint32_t *s_start = reinterpret_cast<uint32_t *>(smem + L.bytes);
    uint32_t *s_pre = s_start + kBMaxLatRows;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&A.ctr[0], 1u);
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= A.ctr[1]) break;   // every workgroup leaves once the queue is drained
        const BItem m = A.items[item];
        const BTile d = A.tiles[m.tile];
        const BReq q = A.reqs[d.req];
        int32_t px0, px1, py0, py1;
        b_tile_rect(q, d.t, px0, px1, py0, py1);
        // lattice rows of the tile (at most the grid's 4096)
        const int32_t nrow = d.gx0 <= d.gx1 ? d.gy1 - d.gy0 + 1 : 0;
        unsigned long long tot = 0;
        for (int32_t k0 = 0; k0 < nrow; k0 += NT) {
            const int32_t k = k0 + (int32_t)threadIdx.x;
            unsigned long long len = 0;
            int64_t st0 = 0;
            if (k < nrow) {
                const int64_t row = (int64_t)(d.gy0 + k) * A.grid;
                st0 = A.cs[row + d.gx0];
                len = (unsigned long long)(A.cs[row + d.gx1 + 1] - st0);
            }
            unsigned long long t2;
            const unsigned long long ex = b_block_exscan(len, sh, t2);
            if (k < nrow) {
                s_start[k] = (uint32_t)st0;
                s_pre[k] = (uint32_t)(tot + ex);
            }
            tot += t2;
        }
        b_region_init(smem, L, threadIdx.x, NT);
        __syncthreads();
        // this part's rows [lo, hi) of the tile's rows
        const unsigned long long E = d.nparts > 1 ? (tot + d.nparts - 1) / d.nparts : tot;
        unsigned long long lo = (unsigned long long)m.part * E, hi = lo + E;
        lo = lo < tot ? lo : tot;
        hi = hi < tot ? hi : tot;
        const int32_t nr = nrow;
        int32_t top = 1;  // binary search over the table: the largest power of two < nr
        while (2 * top < nr) top *= 2;
        // the last lattice row starting at or before j
        auto row_of = [&](unsigned long long j) {
            int32_t k = 0;
            for (int32_t step = top; step >= 1; step >>= 1)
                if (k + step < nr && (unsigned long long)s_pre[k + step] <= j) k += step;
            return k;
        };
        // Each thread keeps a run of rows that land in one cell in registers and adds it to
        // the LDS region when the cell changes: the rows are sorted by lattice cell, so in
        // a dense view a thread's rows (NT apart) fall in one pixel for long stretches, and
        // the run replaces per-row LDS atomics that would all hit the same address.
        uint32_t run_lc = 0, run_n = 0, run_nan = 0;
        double run_s = 0.0;
        auto flush = [&]() {
            if (run_n == 0u) return;
            atomicAdd(reinterpret_cast<uint32_t *>(smem + L.o_cnt) + run_lc, run_n);
            if (L.has_val) {
                if (run_nan) atomicAdd(reinterpret_cast<uint32_t *>(smem + L.o_nan) + run_lc, run_nan);
                if (run_n > run_nan) atomicAdd(reinterpret_cast<double *>(smem + L.o_sum) + run_lc, run_s);
            }
            run_n = 0u;
            run_nan = 0u;
            run_s = 0.0;
        };
        // project kBItems loaded rows and add the tile's ones to the thread's run
        auto bin = [&](const float *xs, const float *ys, const float *vs, const bool *ok) {
#pragma unroll
            for (int it = 0; it < kBItems; ++it) {
                if (!ok[it]) continue;
                const double x = (double)xs[it], y = (double)ys[it];
                if (!((q.xmin <= x && x <= q.xmax) && (q.ymin <= y && y <= q.ymax))) continue;
                int32_t xx = proj(x, q.sx, q.tx), yy = proj(y, q.sy, q.ty);
                xx = xx >= q.W ? q.W - 1 : xx;
                yy = yy >= q.H ? q.H - 1 : yy;
                if (xx < 0) xx += q.W;  // numba's negative-index wraparound
                if (yy < 0) yy += q.H;
                if (xx < px0 || xx >= px1 || yy < py0 || yy >= py1) continue;  // another tile's
                const uint32_t lc = (uint32_t)((yy - py0) * q.tw + (xx - px0));
                if (lc != run_lc) {
                    flush();
                    run_lc = lc;
                }
                ++run_n;
                if (L.has_val) {
                    const double v = (double)vs[it];
                    if (v != v) ++run_nan;
                    else run_s += v;
                }
            }
        };
        // Rows [lo, hi) of the tile's concatenated lattice rows, NT * kBItems a chunk, the
        // next chunk's loads issued before this chunk is binned.  A chunk that lies inside
        // one lattice row (the common case for a wide view, whose lattice rows are long)
        // is one contiguous run of sorted rows, found by one search for the whole chunk;
        // any other chunk finds each row's lattice row by a search of the table.
        struct Set {
            float x[kBItems], y[kBItems], v[kBItems];
            bool ok[kBItems];
        };
        auto load = [&](unsigned long long c, Set &S) {
            const int32_t kc = row_of(c < tot ? c : (tot > 0 ? tot - 1 : 0));  // uniform
            const unsigned long long kend = kc + 1 < nr ? (unsigned long long)s_pre[kc + 1] : tot;
            const bool one = c + (unsigned long long)(NT * kBItems) <= kend;   // uniform
#pragma unroll
            for (int it = 0; it < kBItems; ++it) {
                const unsigned long long j = c + (unsigned long long)(it * NT + threadIdx.x);
                S.ok[it] = j < hi;
                int64_t row;
                if (one) {  // j < kend: a row of lattice row kc (read even past hi, dropped)
                    row = (int64_t)s_start[kc] + (int64_t)(j - s_pre[kc]);
                } else {
                    const unsigned long long jj = S.ok[it] ? j : lo;  // past hi: row lo, dropped
                    const int32_t k = row_of(jj);
                    row = (int64_t)s_start[k] + (int64_t)(jj - s_pre[k]);
                }
                S.x[it] = A.x[row];
                S.y[it] = A.y[row];
                S.v[it] = L.has_val ? A.v[row] : 0.0f;
            }
        };
        if (lo < hi) {
            constexpr unsigned long long kStep = (unsigned long long)NT * kBItems;
            Set sa, sb;
            load(lo, sa);
            // no exit between the two halves: a step past hi loads row lo and bins nothing
            for (unsigned long long c = lo; c < hi; c += 2 * kStep) {
                load(c + kStep, sb);
                bin(sa.x, sa.y, sa.v, sa.ok);
                load(c + 2 * kStep, sa);
                bin(sb.x, sb.y, sb.v, sb.ok);
            }
        }
        flush();
        __syncthreads();
        if (d.nparts > 1) {
            // a split tile's part: the tile's cells of its region go to scratch (same
            // layout), the merge kernel folds them
            if (d.region + m.part < (uint32_t)A.max_regions) {
                uint8_t *dst = A.scratch + (size_t)(d.region + m.part) * L.bytes;
                const uint32_t used = (uint32_t)(q.tw * q.th);
                for (uint32_t c = threadIdx.x; c < used; c += NT) {
                    reinterpret_cast<uint32_t *>(dst + L.o_cnt)[c] = reinterpret_cast<const uint32_t *>(smem + L.o_cnt)[c];
                    if (L.has_val) {
                        reinterpret_cast<uint32_t *>(dst + L.o_nan)[c] = reinterpret_cast<const uint32_t *>(smem + L.o_nan)[c];
                        reinterpret_cast<double *>(dst + L.o_sum)[c] = reinterpret_cast<const double *>(smem + L.o_sum)[c];
                    }
                }
            }
        } else {
            b_store_tile(smem, L, q, d.t, NT);
        }
        __syncthreads();  // the region and table are rebuilt for the next item
    }
}

// ---- merge: split tiles ---------------------------------------------------------------
constexpr int kBMergeThreads = 512;

__global__ __launch_bounds__(kBMergeThreads) void b_merge_kernel(BArgs A) {
    extern __shared__ __align__(16) uint8_t smem[];
    const BLay &L = A.L;
    const uint32_t nsplit = A.ctr[2];
    for (uint32_t s = blockIdx.x; s < nsplit && s < (uint32_t)A.max_tiles; s += gridDim.x) {
        const BTile d = A.tiles[A.split[s]];
        const BReq q = A.reqs[d.req];
        const uint32_t np = d.region + d.nparts <= (uint32_t)A.max_regions ? d.nparts : 0u;
        const uint8_t *parts = A.scratch + (size_t)d.region * L.bytes;
        const uint32_t used = (uint32_t)(q.tw * q.th);  // the tile's cells of each region
        constexpr uint32_t kU = 4;                     // parts whose loads are in flight together
        for (uint32_t c = threadIdx.x; c < used; c += kBMergeThreads) {
            uint32_t n = 0, nn = 0;
            double sum = 0.0;
            uint32_t p = 0;
            for (; p + kU <= np; p += kU) {
                uint32_t cn[kU], cnn[kU];
                double cs[kU];
#pragma unroll
                for (uint32_t u = 0; u < kU; ++u) {
                    const uint8_t *r = parts + (size_t)(p + u) * L.bytes;
                    cn[u] = reinterpret_cast<const uint32_t *>(r + L.o_cnt)[c];
                    cnn[u] = L.has_val ? reinterpret_cast<const uint32_t *>(r + L.o_nan)[c] : 0u;
                    cs[u] = L.has_val ? reinterpret_cast<const double *>(r + L.o_sum)[c] : 0.0;
                }
#pragma unroll
                for (uint32_t u = 0; u < kU; ++u) {
                    n += cn[u];
                    nn += cnn[u];
                    sum += cs[u];
                }
            }
            for (; p < np; ++p) {
                const uint8_t *r = parts + (size_t)p * L.bytes;
                n += reinterpret_cast<const uint32_t *>(r + L.o_cnt)[c];
                if (L.has_val) {
                    nn += reinterpret_cast<const uint32_t *>(r + L.o_nan)[c];
                    sum += reinterpret_cast<const double *>(r + L.o_sum)[c];
                }
            }
            reinterpret_cast<uint32_t *>(smem + L.o_cnt)[c] = n;
            if (L.has_val) {
                reinterpret_cast<uint32_t *>(smem + L.o_nan)[c] = nn;
                reinterpret_cast<double *>(smem + L.o_sum)[c] = sum;
            }
        }
        __syncthreads();
        b_store_tile(smem, L, q, d.t, kBMergeThreads);
        __syncthreads();
    }
}

// ---- host side --------------------------------------------------------------------------
static BLay b_layout(bool has_val, uint32_t cap) {
    BLay L;
    L.has_val = has_val;
    L.cap = cap;
    uint32_t o = 0;
    L.o_sum = o; o += has_val ? 8u * cap : 0u;
    L.o_cnt = o; o += 4u * cap;
    L.o_nan = o; o += has_val ? 4u * cap : 0u;
    L.bytes = (o + 15u) & ~15u;
    L.o_sum = has_val ? L.o_sum : 0u;
    return L;
}

static size_t b_align(size_t x) { return (x + 255) / 256 * 256; }

struct BLayout {
    BLay L;
    size_t o_table, table_bytes, o_reqs, o_tiles, o_treq, o_items, o_split, o_ctr, o_scratch,
        total;
    int32_t max_items, max_regions;
};

static int b_plan(const dsx_batch_config *c, BLayout &P) {
    if (!c) return DSX_ERR_NULL;
    if (c->max_requests < 1 || c->max_tiles < 1 || c->max_regions < 1) return DSX_ERR_SHAPE;
    if (c->grid < 1 || c->grid > kBMaxLatRows || c->n < 0) return DSX_ERR_SHAPE;
    if (c->n >= (1ll << 32)) return DSX_ERR_SHAPE;  // the tile tables hold 32-bit rows
    if (!(c->x1 > c->x0) || !(c->y1 > c->y0)) return DSX_ERR_SHAPE;
    if (c->planes == 0 || (c->planes & ~(uint32_t)(BP_COUNT | BP_COUNTV | BP_SUM | BP_MEAN)))
        return DSX_ERR_MODE;
    const bool has_val = (c->planes & (BP_COUNTV | BP_SUM | BP_MEAN)) != 0;
    if (has_val && !c->val && c->n > 0) return DSX_ERR_ACCUM;
    if (c->n > 0 && (!c->x || !c->y || !c->cell_start)) return DSX_ERR_NULL;
    if (c->threads != 256 && c->threads != 512 && c->threads != 1024) return DSX_ERR_SHAPE;
    if (c->workgroups < 1) return DSX_ERR_SHAPE;
    const uint32_t cb = has_val ? 16u : 4u;
    const uint32_t cap = (uint32_t)((kBLds - kBTable - 1024) / cb) & ~3u;
    P.L = b_layout(has_val, cap);
    P.max_regions = c->max_regions;
    P.max_items = c->max_tiles + c->max_regions;
    size_t o = 0;
    P.table_bytes = (size_t)kBHeader + sizeof(dsx_batch_req) * (size_t)c->max_requests;
    P.o_table = o; o = b_align(o + P.table_bytes);
    P.o_reqs = o; o = b_align(o + sizeof(BReq) * (size_t)c->max_requests);
    P.o_tiles = o; o = b_align(o + sizeof(BTile) * (size_t)c->max_tiles);
    P.o_treq = o; o = b_align(o + 4 * (size_t)c->max_tiles);
    P.o_items = o; o = b_align(o + sizeof(BItem) * (size_t)P.max_items);
    P.o_split = o; o = b_align(o + 4 * (size_t)c->max_tiles);
    P.o_ctr = o; o = b_align(o + 64);
    P.o_scratch = o; o = b_align(o + (size_t)P.L.bytes * (size_t)c->max_regions);
    P.total = o;
    return DSX_OK;
}

}  // namespace dsx

using namespace dsx;

struct dsx_batch {
    dsx_batch_config cfg;
    BLayout P;
    BArgs A;
    void *htable;              // the caller's pinned table
    void *dtable;              // its copy in the workspace
    dsx_batch_req *table;      // host view of the pinned table (after the header)
    BHeader *hdr;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int use_graph;
};

static int b_launch_all(dsx_batch *b, hipStream_t st) {
    const BArgs &A = b->A;
    const dsx_batch_config &c = b->cfg;
    // the request table: pinned host memory -> workspace (a copy node of the graph), so
    // that no kernel reads host memory
    const hipError_t ec = hipMemcpyAsync(b->dtable, b->htable, b->P.table_bytes,
                                         hipMemcpyHostToDevice, st);
    if (ec != hipSuccess) { record_hip_error(ec); return DSX_ERR_HIP; }
    hipLaunchKernelGGL(b_stage_kernel, dim3(c.max_requests < 256 ? c.max_requests : 256), dim3(256),
                       0, st, A);
    const int tg = c.max_tiles < 4096 ? c.max_tiles : 4096;
    hipLaunchKernelGGL(b_tiles_kernel, dim3(tg), dim3(kBTileThreads), 0, st, A);
    hipLaunchKernelGGL(b_items_kernel, dim3(1), dim3(kBItemThreads), 0, st, A);
    const size_t lds = (size_t)b->P.L.bytes + kBTable;
    switch (c.threads) {
    case 256: hipLaunchKernelGGL(b_agg_kernel<256>, dim3(c.workgroups), dim3(256), lds, st, A); break;
    case 512: hipLaunchKernelGGL(b_agg_kernel<512>, dim3(c.workgroups), dim3(512), lds, st, A); break;
    default: hipLaunchKernelGGL(b_agg_kernel<1024>, dim3(c.workgroups), dim3(1024), lds, st, A); break;
    }
    const int mg = c.max_tiles < 1024 ? c.max_tiles : 1024;
    hipLaunchKernelGGL(b_merge_kernel, dim3(mg), dim3(kBMergeThreads), (size_t)b->P.L.bytes, st, A);
    return launch_status();
}

extern "C" size_t dsx_batch_workspace_size(const dsx_batch_config *c) {
    BLayout P;
    if (b_plan(c, P)) return 0;
    return P.total;
}

extern "C" size_t dsx_batch_table_bytes(int32_t max_requests) {
    if (max_requests < 1) return 0;
    return (size_t)kBHeader + sizeof(dsx_batch_req) * (size_t)max_requests;
}

extern "C" int32_t dsx_batch_tile_cells(const dsx_batch_config *c) {
    BLayout P;
    if (b_plan(c, P)) return 0;
    return (int32_t)P.L.cap;
}

extern "C" int dsx_batch_create(const dsx_batch_config *c, void *workspace, size_t workspace_bytes,
                                void *host_table, int32_t use_graph, dsx_batch **out) {
    if (!out) return DSX_ERR_NULL;
    *out = nullptr;
    BLayout P;
    int st = b_plan(c, P);
    if (st) return st;
    if (!workspace || workspace_bytes < P.total) return DSX_ERR_WORKSPACE;
    if (!host_table) return DSX_ERR_NULL;
    uint8_t *dtable = reinterpret_cast<uint8_t *>(workspace) + P.o_table;
    hipError_t e = hipSuccess;
    dsx_batch *b = new (std::nothrow) dsx_batch();
    if (!b) return DSX_ERR_HIP;
    b->cfg = *c;
    b->P = P;
    b->hdr = reinterpret_cast<BHeader *>(host_table);
    b->table = reinterpret_cast<dsx_batch_req *>(reinterpret_cast<uint8_t *>(host_table) + kBHeader);
    memset(host_table, 0, dsx_batch_table_bytes(c->max_requests));
    uint8_t *ws = reinterpret_cast<uint8_t *>(workspace);
    BArgs &A = b->A;
    b->htable = host_table;
    b->dtable = dtable;
    A.hhdr = reinterpret_cast<const BHeader *>(dtable);
    A.host = reinterpret_cast<const dsx_batch_req *>(dtable + kBHeader);
    A.reqs = reinterpret_cast<BReq *>(ws + P.o_reqs);
    A.tiles = reinterpret_cast<BTile *>(ws + P.o_tiles);
    A.tile_req = reinterpret_cast<int32_t *>(ws + P.o_treq);
    A.items = reinterpret_cast<BItem *>(ws + P.o_items);
    A.split = reinterpret_cast<uint32_t *>(ws + P.o_split);
    A.ctr = reinterpret_cast<uint32_t *>(ws + P.o_ctr);
    A.scratch = ws + P.o_scratch;
    A.max_req = c->max_requests;
    A.max_tiles = c->max_tiles;
    A.max_items = P.max_items;
    A.max_regions = P.max_regions;
    A.L = P.L;
    A.n = c->n;
    A.x = c->x; A.y = c->y; A.v = c->val;
    if (c->n == 0) { A.x = A.y = A.v = reinterpret_cast<const float *>(c->cell_start); }
    A.grid = c->grid;
    A.ex0 = c->x0; A.esx = (double)c->grid / (c->x1 - c->x0);
    A.ey0 = c->y0; A.esy = (double)c->grid / (c->y1 - c->y0);
    A.cs = c->cell_start;
    A.planes = c->planes;
    const int lds = (int)(P.L.bytes + kBTable);
    const void *kerns[3] = {(const void *)b_agg_kernel<256>, (const void *)b_agg_kernel<512>,
                            (const void *)b_agg_kernel<1024>};
    for (const void *k : kerns) {
        e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) { record_hip_error(e); delete b; return DSX_ERR_HIP; }
    }
    e = hipFuncSetAttribute((const void *)b_merge_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)P.L.bytes);
    if (e != hipSuccess) { record_hip_error(e); delete b; return DSX_ERR_HIP; }
    b->use_graph = use_graph ? 1 : 0;
    b->graph = nullptr;
    b->exec = nullptr;
    if (b->use_graph) {
        // capture the kernel sequence once on a private stream; every batch replays it
        hipStream_t cs = nullptr;
        e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
        if (e != hipSuccess) { record_hip_error(e); delete b; return DSX_ERR_HIP; }
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            st = b_launch_all(b, cs);
            hipGraph_t g = nullptr;
            const hipError_t e2 = hipStreamEndCapture(cs, &g);
            if (st == DSX_OK && e2 != hipSuccess) { record_hip_error(e2); st = DSX_ERR_HIP; }
            if (st == DSX_OK) {
                b->graph = g;
                e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
                if (e != hipSuccess) { record_hip_error(e); st = DSX_ERR_HIP; }
            } else if (g) {
                (void)hipGraphDestroy(g);
            }
        } else {
            record_hip_error(e);
            st = DSX_ERR_HIP;
        }
        (void)hipStreamDestroy(cs);
        if (st) {
            if (b->exec) (void)hipGraphExecDestroy(b->exec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            delete b;
            return st;
        }
    }
    *out = b;
    return DSX_OK;
}

// Validate the nreq requests the caller wrote into the pinned table, fix each request's
// tile shape (at most the LDS tile's cells) and tile base, write the header and launch
// the batch on `stream`.
extern "C" int dsx_batch_submit(dsx_batch *b, int32_t nreq, void *stream) {
    if (!b) return DSX_ERR_NULL;
    const dsx_batch_config &c = b->cfg;
    if (nreq < 0 || nreq > c.max_requests) return DSX_ERR_SHAPE;
    const int64_t cap = (int64_t)b->P.L.cap;
    const bool has_val = b->P.L.has_val;
    int64_t tiles = 0;
    for (int32_t r = 0; r < nreq; ++r) {
        dsx_batch_req &q = b->table[r];
        if (q.width < 1 || q.height < 1 || q.width >= (1 << 24) || q.height >= (1 << 24))
            return DSX_ERR_SHAPE;
        if (!(q.sx > 0) || !(q.sy > 0)) return DSX_ERR_SHAPE;
        if (((c.planes & BP_COUNT) && !q.count) || ((c.planes & BP_COUNTV) && !q.countv) ||
            ((c.planes & BP_SUM) && !q.sum) || ((c.planes & BP_MEAN) && !q.mean))
            return DSX_ERR_NULL;
        if (!(c.planes & BP_COUNT)) q.count = nullptr;
        if (!has_val || !(c.planes & BP_COUNTV)) q.countv = nullptr;
        if (!has_val || !(c.planes & BP_SUM)) q.sum = nullptr;
        if (!has_val || !(c.planes & BP_MEAN)) q.mean = nullptr;
        int64_t th = q.tile_h > 0 ? q.tile_h : 64;
        int64_t tw = q.tile_w > 0 ? q.tile_w : cap / th;
        if (th > q.height) th = q.height;
        if (th > cap) th = cap;
        if (tw * th > cap) tw = cap / th;
        if (tw > q.width) tw = q.width;
        if (tw < 1) tw = 1;
        q.tile_w = (int32_t)tw;
        q.tile_h = (int32_t)th;
        if (q.part_rows < 1) q.part_rows = 1 << 16;
        q.tile_base = (int32_t)tiles;
        tiles += ((q.width + tw - 1) / tw) * ((q.height + th - 1) / th);
        if (tiles > c.max_tiles) return DSX_ERR_SHAPE;
    }
    b->hdr->nreq = nreq;
    b->hdr->ntiles = (int32_t)tiles;
    if (nreq == 0) return DSX_OK;
    hipStream_t st = as_stream(stream);
    if (b->use_graph) {
        const hipError_t e = hipGraphLaunch(b->exec, st);
        if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
        return DSX_OK;
    }
    return b_launch_all(b, st);
}

extern "C" dsx_batch_req *dsx_batch_table(dsx_batch *b) { return b ? b->table : nullptr; }

extern "C" int dsx_batch_destroy(dsx_batch *b) {
    if (!b) return DSX_OK;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    delete b;
    return DSX_OK;
}
