// dsx_serve.hip -- grid index and row-range gather for a resident point set.
//
// A viewport server answers many Canvas.points calls over one resident point set.  The
// index sorts the rows once by the cell of a grid x grid lattice over the set's extent
// (rocPRIM radix sort of (cell, row) pairs, rows moved as packed structures), so that a
// viewport gathers only the rows of the cells it overlaps and aggregates those with
// dsx_points_agg.  The index changes which rows a call reads, never which rows land in
// the canvas: the viewport's own range test still decides.
#include <string.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "dsx_common.h"

namespace dsx {

template <typename T, int NC>
struct PackedRow {
    T c[NC];
};

// Cell of one row; grid*grid for a row with a NaN coordinate (sorted last, in no cell).
template <typename T>
__device__ __forceinline__ uint32_t grid_cell(T xv, T yv, double x0, double sx, double y0,
                                              double sy, int32_t grid) {
    const double x = (double)xv, y = (double)yv;
    if (x != x || y != y) return (uint32_t)grid * (uint32_t)grid;
    // rows outside the extent (infinities included) go to the edge cells
    const double top = (double)(grid - 1);
    double fx = (x - x0) * sx, fy = (y - y0) * sy;
    fx = fx < 0.0 ? 0.0 : (fx > top ? top : fx);
    fy = fy < 0.0 ? 0.0 : (fy > top ? top : fy);
    return (uint32_t)((int64_t)fy * grid + (int64_t)fx);
}

template <typename T, int NC>
__global__ __launch_bounds__(256) void grid_pack_kernel(int64_t n, const T *x, const T *y,
                                                        const T *v, double x0, double sx,
                                                        double y0, double sy, int32_t grid,
                                                        uint32_t *keys, PackedRow<T, NC> *rows) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        PackedRow<T, NC> r;
        r.c[0] = x[i];
        r.c[1] = y[i];
        if constexpr (NC == 3) r.c[2] = v[i];
        keys[i] = grid_cell<T>(r.c[0], r.c[1], x0, sx, y0, sy, grid);
        rows[i] = r;
    }
}

template <typename T, int NC>
__global__ __launch_bounds__(256) void grid_unpack_kernel(int64_t n, const PackedRow<T, NC> *rows,
                                                          T *x, T *y, T *v) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const PackedRow<T, NC> r = rows[i];
        x[i] = r.c[0];
        y[i] = r.c[1];
        if constexpr (NC == 3) v[i] = r.c[2];
    }
}

// cell_start[c] = first sorted row whose key is >= c (binary search), c in [0, cells].
__global__ __launch_bounds__(256) void cell_start_kernel(int64_t n, const uint32_t *keys,
                                                         int64_t cells, int64_t *cell_start) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c <= cells; c += stride) {
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = lo + ((hi - lo) >> 1);
            if ((int64_t)keys[mid] < c) lo = mid + 1;
            else hi = mid;
        }
        cell_start[c] = lo;
    }
}

struct IndexPlan {
    size_t o_keys_in, o_keys_out, o_rows_in, o_rows_out, o_temp, temp_bytes, total;
    int end_bit;
};

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

template <typename T, int NC>
static int index_plan(const dsx_grid_index_desc *g, IndexPlan &P) {
    const int64_t n = g->n;
    const uint64_t cells = (uint64_t)g->grid * (uint64_t)g->grid;
    P.end_bit = 1;
    while ((1ull << P.end_bit) <= cells) ++P.end_bit;  // keys go up to `cells` (NaN rows)
    size_t o = 0;
    P.o_keys_in = o; o = align256(o + 4 * (size_t)n);
    P.o_keys_out = o; o = align256(o + 4 * (size_t)n);
    P.o_rows_in = o; o = align256(o + sizeof(PackedRow<T, NC>) * (size_t)n);
    P.o_rows_out = o; o = align256(o + sizeof(PackedRow<T, NC>) * (size_t)n);
    P.temp_bytes = 0;
    const hipError_t e = rocprim::radix_sort_pairs(
        nullptr, P.temp_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
        (const PackedRow<T, NC> *)nullptr, (PackedRow<T, NC> *)nullptr, (size_t)n, 0,
        P.end_bit, (hipStream_t)0);
    if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
    P.o_temp = o; o = align256(o + P.temp_bytes);
    P.total = o;
    return DSX_OK;
}

template <typename T, int NC>
static int index_build(const dsx_grid_index_desc *g, uint8_t *ws, size_t ws_bytes,
                       hipStream_t st) {
    IndexPlan P;
    int s = index_plan<T, NC>(g, P);
    if (s) return s;
    if (ws_bytes < P.total) return DSX_ERR_WORKSPACE;
    uint32_t *keys_in = reinterpret_cast<uint32_t *>(ws + P.o_keys_in);
    uint32_t *keys_out = reinterpret_cast<uint32_t *>(ws + P.o_keys_out);
    auto *rows_in = reinterpret_cast<PackedRow<T, NC> *>(ws + P.o_rows_in);
    auto *rows_out = reinterpret_cast<PackedRow<T, NC> *>(ws + P.o_rows_out);
    const double sx = (double)g->grid / (g->x1 - g->x0);
    const double sy = (double)g->grid / (g->y1 - g->y0);
    const int grid = grid_for(g->n, 256);
    hipLaunchKernelGGL((grid_pack_kernel<T, NC>), dim3(grid), dim3(256), 0, st, g->n,
                       (const T *)g->x, (const T *)g->y, (const T *)g->val, g->x0, sx, g->y0,
                       sy, g->grid, keys_in, rows_in);
    if ((s = launch_status())) return s;
    size_t temp_bytes = P.temp_bytes;
    hipError_t e = rocprim::radix_sort_pairs(ws + P.o_temp, temp_bytes, keys_in, keys_out,
                                             rows_in, rows_out, (size_t)g->n, 0, P.end_bit, st);
    if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
    hipLaunchKernelGGL((grid_unpack_kernel<T, NC>), dim3(grid), dim3(256), 0, st, g->n,
                       (const PackedRow<T, NC> *)rows_out, (T *)g->out_x, (T *)g->out_y,
                       (T *)g->out_val);
    if ((s = launch_status())) return s;
    const int64_t cells = (int64_t)g->grid * g->grid;
    hipLaunchKernelGGL(cell_start_kernel, dim3(grid_for(cells + 1, 256)), dim3(256), 0, st, g->n,
                       (const uint32_t *)keys_out, cells, g->cell_start);
    return launch_status();
}

static int check_index(const dsx_grid_index_desc *g) {
    if (!g) return DSX_ERR_NULL;
    if (g->n < 1 || g->grid < 1 || g->grid > 4096) return DSX_ERR_SHAPE;
    if (g->dtype != DSX_F32 && g->dtype != DSX_F64) return DSX_ERR_DTYPE;
    if (!g->x || !g->y || !g->out_x || !g->out_y || !g->cell_start) return DSX_ERR_NULL;
    if ((g->val == nullptr) != (g->out_val == nullptr)) return DSX_ERR_NULL;
    if (!(g->x1 > g->x0) || !(g->y1 > g->y0)) return DSX_ERR_SHAPE;
    return DSX_OK;
}

// One workgroup per item: rows [src, src + len) of each column to [dst, dst + len).
__global__ __launch_bounds__(256) void gather_rows_kernel(dsx_gather_desc g) {
    const int64_t k = blockIdx.x;
    const int64_t src = g.items[3 * k], dst = g.items[3 * k + 1], len = g.items[3 * k + 2];
    for (int c = 0; c < g.ncols; ++c) {
        if (g.elem_bytes[c] == 4) {
            const uint32_t *s = reinterpret_cast<const uint32_t *>(g.src[c]) + src;
            uint32_t *d = reinterpret_cast<uint32_t *>(g.dst[c]) + dst;
            for (int64_t i = threadIdx.x; i < len; i += blockDim.x) d[i] = s[i];
        } else {
            const uint64_t *s = reinterpret_cast<const uint64_t *>(g.src[c]) + src;
            uint64_t *d = reinterpret_cast<uint64_t *>(g.dst[c]) + dst;
            for (int64_t i = threadIdx.x; i < len; i += blockDim.x) d[i] = s[i];
        }
    }
}

}  // namespace dsx

using namespace dsx;

extern "C" size_t dsx_grid_index_workspace_size(const dsx_grid_index_desc *g) {
    if (check_index(g)) return 0;
    IndexPlan P;
    int s;
    if (g->dtype == DSX_F32) s = g->val ? index_plan<float, 3>(g, P) : index_plan<float, 2>(g, P);
    else s = g->val ? index_plan<double, 3>(g, P) : index_plan<double, 2>(g, P);
    return s ? 0 : P.total;
}

extern "C" int dsx_grid_index_build(const dsx_grid_index_desc *g, void *workspace,
                                    size_t workspace_bytes, void *stream) {
    const int s = check_index(g);
    if (s) return s;
    if (!workspace) return DSX_ERR_WORKSPACE;
    uint8_t *ws = reinterpret_cast<uint8_t *>(workspace);
    hipStream_t st = as_stream(stream);
    if (g->dtype == DSX_F32)
        return g->val ? index_build<float, 3>(g, ws, workspace_bytes, st)
                      : index_build<float, 2>(g, ws, workspace_bytes, st);
    return g->val ? index_build<double, 3>(g, ws, workspace_bytes, st)
                  : index_build<double, 2>(g, ws, workspace_bytes, st);
}

extern "C" int dsx_gather_rows(const dsx_gather_desc *g, void *stream) {
    if (!g) return DSX_ERR_NULL;
    if (g->ncols < 1 || g->ncols > 4 || g->nitems < 0 || g->nitems > 0x7FFFFFFF)
        return DSX_ERR_SHAPE;
    if (g->nitems == 0) return DSX_OK;
    if (!g->items) return DSX_ERR_NULL;
    for (int c = 0; c < g->ncols; ++c) {
        if (!g->src[c] || !g->dst[c]) return DSX_ERR_NULL;
        if (g->elem_bytes[c] != 4 && g->elem_bytes[c] != 8) return DSX_ERR_DTYPE;
    }
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)g->nitems), dim3(256), 0,
                       as_stream(stream), *g);
    return launch_status();
}
