// dsx_nselect.hip -- the n-value reductions of the points path: max_n, min_n, first_n,
// last_n and their row-index forms (reductions.py:1457-1805, 2375-2562).
//
// The reference inserts every row into a per-pixel sorted list of n slots.  Here, for
// n <= kNpassMax (float32 values for max_n / min_n), n streaming passes each keep one
// slot of every cell: a row's packed 64-bit order key is distinct per row and orders as
// the list does, and pass k keeps the best key past the one pass k - 1 kept.  Otherwise
// the kept rows are sorted once on the device by (cell, order key) with rocPRIM's stable
// radix sort -- rows that tie on the order key keep their row order, which is the order
// the reference's strict comparisons leave them in -- and the first n rows of every
// cell's run are written out.
#include <string.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "dsx_project.h"

namespace dsx {
namespace {

constexpr uint32_t kDropped = 0xFFFFFFFFu;  // cell of a row that is not kept

// Cell of row i, or kDropped (outside the view, category dropped, NaN value when the
// value is checked).
template <typename TC, typename TV>
__device__ __forceinline__ uint32_t cell_of(const PointArgs &a, int64_t i, int nan_check) {
    int32_t px, py, pc;
    if (!locate<TC>(a, i, stream_load(reinterpret_cast<const TC *>(a.x) + i),
                    stream_load(reinterpret_cast<const TC *>(a.y) + i), px, py, pc))
        return kDropped;
    if (nan_check) {
        const TV v = reinterpret_cast<const TV *>(a.val)[i];
        if (v != v) return kDropped;
    }
    return (uint32_t)(((int64_t)py * a.W + px) * a.cat.ncat + pc);
}

// Order key of a value: ascending keys sort min_n's order; their complement max_n's.
template <typename TV>
__device__ __forceinline__ uint64_t value_key(TV v, int descending) {
    uint64_t k;
    if constexpr (sizeof(TV) == 4) k = fkey32((float)v);
    else k = fkey64((double)v);
    if (descending) k = (sizeof(TV) == 4 ? 0xFFFFFFFFull : ~0ull) - k;
    return k;
}

// Row-ordered modes and 32-bit value keys: one composite key (cell << 32 | order).  A
// dropped row takes cell `ncells`, so the keys span 32 + bit_width(ncells) bits and the
// sort runs only the passes those bits need.
template <typename TC, typename TV>
__global__ __launch_bounds__(256) void composite_kernel(PointArgs a, int mode, int nan_check,
                                                        uint32_t ncells, uint64_t *keys,
                                                        uint32_t *rows) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        uint32_t c = cell_of<TC, TV>(a, i, nan_check);
        uint64_t order;
        if (mode == DSX_NS_FIRST) order = (uint64_t)i;
        else if (mode == DSX_NS_LAST) order = 0xFFFFFFFFull - (uint64_t)i;
        else order = c == kDropped ? 0ull : value_key<TV>(reinterpret_cast<const TV *>(a.val)[i],
                                                          mode == DSX_NS_MAX);
        if (c == kDropped) c = ncells;
        keys[i] = ((uint64_t)c << 32) | order;
        rows[i] = (uint32_t)i;
    }
}

// 64-bit value keys, first sort: by value order alone (rows start in row order).
template <typename TV>
__global__ __launch_bounds__(256) void value_key_kernel(PointArgs a, int descending,
                                                        uint64_t *keys, uint32_t *rows) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        const TV v = reinterpret_cast<const TV *>(a.val)[i];
        keys[i] = v != v ? 0ull : value_key<TV>(v, descending);
        rows[i] = (uint32_t)i;
    }
}

// Second sort's keys: the cell of each row in value order (`ncells` when dropped).
template <typename TC, typename TV>
__global__ __launch_bounds__(256) void cell_key_kernel(PointArgs a, const uint32_t *rows,
                                                       uint32_t ncells, uint32_t *cells) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < a.n; j += stride) {
        const uint32_t c = cell_of<TC, TV>(a, rows[j], 1);
        cells[j] = c == kDropped ? ncells : c;
    }
}

// The first n rows of every cell's run of the sorted order.
__global__ __launch_bounds__(256) void take_runs_kernel(int64_t n, const uint64_t *keys64,
                                                        const uint32_t *keys32, int shift,
                                                        const uint32_t *rows, int nsel,
                                                        uint32_t ncells, int64_t *index) {
    auto cell = [&](int64_t j) -> uint32_t {
        return keys64 ? (uint32_t)(keys64[j] >> shift) : keys32[j];
    };
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const uint32_t c = cell(j);
        if (c >= ncells) continue;  // dropped rows
        if (j > 0 && cell(j - 1) == c) continue;  // not the start of the cell's run
        for (int k = 0; k < nsel && j + k < n; ++k) {
            if (cell(j + k) != c) break;
            index[(int64_t)c * nsel + k] = (int64_t)rows[j + k];
        }
    }
}

// ---- the n-pass form --------------------------------------------------------------------
// ri_npass (the row-index window pass over packed 64-bit keys) keeps slot k of every cell
// in pass k; index[cell * n + k] holds the slot's packed key -- MAX key(v) << 32 | ~row,
// MIN key(v) << 32 | row, FIRST row, LAST ~row (32-bit); 0 (MAX) or all ones: empty --
// until this turns it into the row.
constexpr int kNpassMax = 8;

__global__ __launch_bounds__(256) void npass_unpack_kernel(int64_t n, int mode, unsigned long long *index) {
    const bool best_max = mode == DSX_NS_MAX;
    const unsigned long long empty = best_max ? 0ull : ~0ull;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long v = index[i];
        const uint32_t lo32 = (uint32_t)v;
        const uint32_t row = (best_max || mode == DSX_NS_LAST) ? ~lo32 : lo32;
        index[i] = v == empty ? ~0ull : (unsigned long long)row;
    }
}

// n <= kNpassMax, and float32 values wherever values are read (max_n / min_n, a NaN check)
bool np_eligible(const dsx_points_desc *d, const dsx_nselect_desc *r) {
    const bool value = r->mode == DSX_NS_MAX || r->mode == DSX_NS_MIN;
    const bool reads = value || r->nan_check;
    return r->n <= kNpassMax && (!reads || (has_value(d) && d->val_dtype == DSX_F32));
}

struct NsPlan {
    bool two_sorts;
    uint32_t ncells;
    int cell_bits;  // bit width of ncells: cells and the dropped marker fit
    size_t o_k0, o_k1, o_r0, o_r1, o_c0, o_c1, o_tmp, tmp_bytes, total;
};

size_t al(size_t x) { return (x + 255) / 256 * 256; }

int ns_plan(const dsx_points_desc *d, const dsx_view *v, const dsx_nselect_desc *r,
            NsPlan &P) {
    const size_t n = (size_t)d->n;
    const bool value = r->mode == DSX_NS_MAX || r->mode == DSX_NS_MIN;
    P.two_sorts = value && d->val_dtype == DSX_F64;
    P.ncells = (uint32_t)((int64_t)v->width * v->height * cat_of(d->cat).ncat);  // < kDropped
    P.cell_bits = 1;
    while (P.cell_bits < 32 && (1ull << P.cell_bits) <= P.ncells) ++P.cell_bits;
    size_t o = 0;
    P.o_k0 = o; o = al(o + 8 * n);
    P.o_k1 = o; o = al(o + 8 * n);
    P.o_r0 = o; o = al(o + 4 * n);
    P.o_r1 = o; o = al(o + 4 * n);
    P.o_c0 = o; o = al(o + (P.two_sorts ? 4 * n : 0));
    P.o_c1 = o; o = al(o + (P.two_sorts ? 4 * n : 0));
    size_t t1 = 0, t2 = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, t1, (const uint64_t *)nullptr,
                                             (uint64_t *)nullptr, (const uint32_t *)nullptr,
                                             (uint32_t *)nullptr, n, 0,
                                             P.two_sorts ? 64 : 32 + P.cell_bits, (hipStream_t)0);
    if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
    if (P.two_sorts) {
        e = rocprim::radix_sort_pairs(nullptr, t2, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                      (const uint32_t *)nullptr, (uint32_t *)nullptr, n, 0,
                                      P.cell_bits, (hipStream_t)0);
        if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
    }
    P.tmp_bytes = t1 > t2 ? t1 : t2;
    P.o_tmp = o; o = al(o + P.tmp_bytes);
    P.total = o;
    return DSX_OK;
}

int ns_check(const dsx_points_desc *d, const dsx_view *v, const dsx_nselect_desc *r) {
    int st = check_points(d, v);
    if (st) return st;
    if (!r) return DSX_ERR_NULL;
    if (r->mode < DSX_NS_FIRST || r->mode > DSX_NS_MIN) return DSX_ERR_MODE;
    if (r->n < 1 || r->n > 4096) return DSX_ERR_SHAPE;
    const bool value = r->mode == DSX_NS_MAX || r->mode == DSX_NS_MIN;
    if ((value || r->nan_check) && !has_value(d)) return DSX_ERR_ACCUM;
    const int64_t cells = (int64_t)v->width * v->height * cat_of(d->cat).ncat;
    if (cells >= (int64_t)kDropped) return DSX_ERR_SHAPE;
    return DSX_OK;
}

template <typename TC, typename TV>
int ns_run(const dsx_points_desc *d, const dsx_view *v, const dsx_nselect_desc *r,
           const NsPlan &P, uint8_t *ws, hipStream_t st) {
    const PointArgs a = args_of(d, v);
    const int64_t n = d->n;
    const int grid = grid_for(n, 256, 8);
    uint64_t *k0 = reinterpret_cast<uint64_t *>(ws + P.o_k0);
    uint64_t *k1 = reinterpret_cast<uint64_t *>(ws + P.o_k1);
    uint32_t *r0 = reinterpret_cast<uint32_t *>(ws + P.o_r0);
    uint32_t *r1 = reinterpret_cast<uint32_t *>(ws + P.o_r1);
    void *tmp = ws + P.o_tmp;
    size_t tb = P.tmp_bytes;
    const bool value = r->mode == DSX_NS_MAX || r->mode == DSX_NS_MIN;
    const int nan_check = (value || r->nan_check) ? 1 : 0;
    int s;
    hipError_t e;
    if (!P.two_sorts) {
        hipLaunchKernelGGL((composite_kernel<TC, TV>), dim3(grid), dim3(256), 0, st, a, r->mode,
                           nan_check, P.ncells, k0, r0);
        if ((s = launch_status())) return s;
        e = rocprim::radix_sort_pairs(tmp, tb, k0, k1, r0, r1, (size_t)n, 0, 32 + P.cell_bits, st);
        if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
        hipLaunchKernelGGL(take_runs_kernel, dim3(grid), dim3(256), 0, st, n,
                           (const uint64_t *)k1, (const uint32_t *)nullptr, 32,
                           (const uint32_t *)r1, r->n, P.ncells, r->index);
        return launch_status();
    }
    uint32_t *c0 = reinterpret_cast<uint32_t *>(ws + P.o_c0);
    uint32_t *c1 = reinterpret_cast<uint32_t *>(ws + P.o_c1);
    hipLaunchKernelGGL((value_key_kernel<TV>), dim3(grid), dim3(256), 0, st, a,
                       (int)(r->mode == DSX_NS_MAX), k0, r0);
    if ((s = launch_status())) return s;
    e = rocprim::radix_sort_pairs(tmp, tb, k0, k1, r0, r1, (size_t)n, 0, 64, st);
    if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
    hipLaunchKernelGGL((cell_key_kernel<TC, TV>), dim3(grid), dim3(256), 0, st, a,
                       (const uint32_t *)r1, P.ncells, c0);
    if ((s = launch_status())) return s;
    tb = P.tmp_bytes;
    e = rocprim::radix_sort_pairs(tmp, tb, c0, c1, r1, r0, (size_t)n, 0, P.cell_bits, st);
    if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
    hipLaunchKernelGGL(take_runs_kernel, dim3(grid), dim3(256), 0, st, n,
                       (const uint64_t *)nullptr, (const uint32_t *)c1, 0, (const uint32_t *)r0,
                       r->n, P.ncells, r->index);
    return launch_status();
}

}  // namespace
}  // namespace dsx

using namespace dsx;

extern "C" size_t dsx_points_nselect_workspace_size(const dsx_points_desc *d, const dsx_view *v,
                                                    const dsx_nselect_desc *r) {
    if (ns_check(d, v, r)) return 0;
    if (np_eligible(d, r)) return ri_npass_workspace_size(d, v);
    NsPlan P;
    if (ns_plan(d, v, r, P)) return 0;
    return P.total;
}

extern "C" int dsx_points_nselect(const dsx_points_desc *d, const dsx_view *v,
                                  const dsx_nselect_desc *r, void *workspace,
                                  size_t workspace_bytes, void *stream) {
    int s = ns_check(d, v, r);
    if (s) return s;
    if (!r->index) return DSX_ERR_NULL;
    hipStream_t st = as_stream(stream);
    const int64_t cells = (int64_t)v->width * v->height * cat_of(d->cat).ncat * r->n;
    if (np_eligible(d, r)) {
        if ((s = fill_async(r->index, DSX_FILL_ONES64, cells, st))) return s;
        if (d->n == 0) return DSX_OK;
        if (r->mode == DSX_NS_MAX && (s = fill_async(r->index, DSX_FILL_ZERO64, cells, st))) return s;
        if ((s = ri_npass(d, v, r->mode, r->n, r->nan_check, r->index, workspace, workspace_bytes, st)))
            return s;
        hipLaunchKernelGGL(npass_unpack_kernel, dim3(grid_for(cells, 256, 4)), dim3(256), 0, st, cells,
                           r->mode, reinterpret_cast<unsigned long long *>(r->index));
        return launch_status();
    }
    NsPlan P;
    if ((s = ns_plan(d, v, r, P))) return s;
    if (!workspace || workspace_bytes < P.total) return DSX_ERR_WORKSPACE;
    if ((s = fill_async(r->index, DSX_FILL_ONES64, cells, st))) return s;
    if (d->n == 0) return DSX_OK;
    uint8_t *ws = reinterpret_cast<uint8_t *>(workspace);
    const int vdt = has_value(d) ? d->val_dtype : DSX_F32;
    if (d->xy_dtype == DSX_F32)
        return vdt == DSX_F32 ? ns_run<float, float>(d, v, r, P, ws, st)
                              : ns_run<float, double>(d, v, r, P, ws, st);
    return vdt == DSX_F32 ? ns_run<double, float>(d, v, r, P, ws, st)
                          : ns_run<double, double>(d, v, r, P, ws, st);
}
