// dsx_misc.hip -- autorange bounds, plane fill, finalize, status strings.
#include <math.h>
#include <string.h>

#include <mutex>

#include "dsx_common.h"

namespace dsx {

static thread_local char g_last_hip[256] = "";

void record_hip_error(hipError_t e) {
    strncpy(g_last_hip, hipGetErrorString(e), sizeof(g_last_hip) - 1);
    g_last_hip[sizeof(g_last_hip) - 1] = 0;
}

int launch_status() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        record_hip_error(e);
        return DSX_ERR_HIP;
    }
    return DSX_OK;
}

int grid_for(int64_t n, int block, int per_cu) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            c > 0)
            cus = c;
    }
    const int64_t want = (n + block - 1) / block;
    const int64_t cap = (int64_t)cus * per_cu;
    return (int)(want < 1 ? 1 : (want < cap ? want : cap));
}

// ---- bounds: glyph.py:66-78 ------------------------------------------------------------
constexpr int kBoundsBlock = 256;

__device__ __forceinline__ void minmax_reduce_block(double &lo, double &hi, double *sh) {
    for (int off = 32; off > 0; off >>= 1) {
        const double olo = __shfl_xor(lo, off, 64);
        const double ohi = __shfl_xor(hi, off, 64);
        lo = olo < lo ? olo : lo;
        hi = ohi > hi ? ohi : hi;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        sh[2 * wave] = lo;
        sh[2 * wave + 1] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            lo = sh[2 * w] < lo ? sh[2 * w] : lo;
            hi = sh[2 * w + 1] > hi ? sh[2 * w + 1] : hi;
        }
    }
}

// kBoundsU rows per thread per iteration, a grid stride apart, all loads issued before
// any is used (one load in flight per thread left the pass at a quarter of HBM speed);
// the column streams past the caches (read once).
constexpr int kBoundsU = 8;

template <typename T>
__global__ __launch_bounds__(kBoundsBlock) void bounds_partial(const T *a, int64_t n,
                                                              double *part) {
    __shared__ double sh[2 * (kBoundsBlock / 64)];
    T lo = (T)INFINITY, hi = (T)-INFINITY;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (kBoundsU - 1) * stride < n; i += kBoundsU * stride) {
        T x[kBoundsU];
#pragma unroll
        for (int u = 0; u < kBoundsU; ++u) x[u] = stream_load(a + i + u * stride);
#pragma unroll
        for (int u = 0; u < kBoundsU; ++u) {
            if (x[u] == x[u]) {
                lo = x[u] < lo ? x[u] : lo;
                hi = x[u] > hi ? x[u] : hi;
            }
        }
    }
    for (; i < n; i += stride) {
        const T x = a[i];
        if (x == x) {
            lo = x < lo ? x : lo;
            hi = x > hi ? x : hi;
        }
    }
    double dlo = (double)lo, dhi = (double)hi;
    minmax_reduce_block(dlo, dhi, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = dlo;
        part[2 * blockIdx.x + 1] = dhi;
    }
}

__global__ __launch_bounds__(kBoundsBlock) void bounds_final(const double *part, int nblk,
                                                            double *out2) {
    __shared__ double sh[2 * (kBoundsBlock / 64)];
    double lo = INFINITY, hi = -INFINITY;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        lo = part[2 * b] < lo ? part[2 * b] : lo;
        hi = part[2 * b + 1] > hi ? part[2 * b + 1] : hi;
    }
    minmax_reduce_block(lo, hi, sh);
    if (threadIdx.x == 0) {
        out2[0] = lo;
        out2[1] = hi;
    }
}

// ---- fill ---------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(T *p, T value, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = value;
}

// ---- finalize -----------------------------------------------------------------------------
// sum-based modes: four cells per thread per iteration, their loads issued together
__device__ __forceinline__ double fin_value(int mode, uint32_t c, double s) {
    switch (mode) {
    case DSX_FIN_SUM: return c > 0u ? s : (double)NAN;
    case DSX_FIN_SUMZERO: return s;
    case DSX_FIN_MEAN: return c > 0u ? s / (double)c : (double)NAN;
    case DSX_FIN_VAR: return c > 0u ? s / (double)c : (double)NAN;
    default: return c > 0u ? sqrt(s / (double)c) : (double)NAN;
    }
}

__global__ __launch_bounds__(256) void finalize_kernel(dsx_finalize_desc f) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const bool f32 = f.key_bytes == 4;
    double *out = reinterpret_cast<double *>(f.out);
    if (f.mode != DSX_FIN_ANY && f.mode != DSX_FIN_MIN && f.mode != DSX_FIN_MAX) {
        constexpr int kU = 4;
        int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        for (; i + (kU - 1) * stride < f.ncells; i += kU * stride) {
            uint32_t c[kU];
            double s[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                c[u] = f.countv[i + u * stride];
                s[u] = f.sum[i + u * stride];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) out[i + u * stride] = fin_value(f.mode, c[u], s[u]);
        }
        for (; i < f.ncells; i += stride) out[i] = fin_value(f.mode, f.countv[i], f.sum[i]);
        return;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < f.ncells; i += stride) {
        switch (f.mode) {
        case DSX_FIN_ANY:
            reinterpret_cast<uint8_t *>(f.out)[i] = f.count[i] > 0u;
            break;
        case DSX_FIN_MIN:
        case DSX_FIN_MAX: {
            double r;
            if (f32) {
                const uint32_t k = reinterpret_cast<const uint32_t *>(f.key)[i];
                const bool empty = f.mode == DSX_FIN_MIN ? k == DSX_MIN_EMPTY32 : k == 0u;
                r = empty ? (double)NAN : (double)unkey32(k);
            } else {
                const uint64_t k = reinterpret_cast<const uint64_t *>(f.key)[i];
                const bool empty = f.mode == DSX_FIN_MIN ? k == DSX_MIN_EMPTY64 : k == 0ull;
                r = empty ? (double)NAN : unkey64(k);
            }
            out[i] = r;
            break;
        }
        default:
            break;  // the sum-based modes took the loop above
        }
    }
}

}  // namespace dsx

using namespace dsx;

extern "C" size_t dsx_bounds_workspace_size(int64_t n) {
    return (size_t)grid_for(n < 1 ? 1 : n, kBoundsBlock, 8) * 2 * sizeof(double);
}

extern "C" int dsx_bounds(const void *col, int32_t dtype, int64_t n, double *out2,
                          void *workspace, size_t workspace_bytes, void *stream) {
    if (!out2 || (n > 0 && !col)) return DSX_ERR_NULL;
    if (n < 0) return DSX_ERR_SHAPE;
    if (dtype != DSX_F32 && dtype != DSX_F64) return DSX_ERR_DTYPE;
    const int grid = grid_for(n < 1 ? 1 : n, kBoundsBlock, 8);
    if (!workspace || workspace_bytes < (size_t)grid * 2 * sizeof(double)) return DSX_ERR_WORKSPACE;
    hipStream_t s = as_stream(stream);
    double *part = reinterpret_cast<double *>(workspace);
    if (dtype == DSX_F32)
        bounds_partial<float><<<grid, kBoundsBlock, 0, s>>>((const float *)col, n, part);
    else
        bounds_partial<double><<<grid, kBoundsBlock, 0, s>>>((const double *)col, n, part);
    int st = launch_status();
    if (st) return st;
    bounds_final<<<1, kBoundsBlock, 0, s>>>(part, grid, out2);
    return launch_status();
}

// ---- bounds without allocation ------------------------------------------------------------
// The per-workgroup partials live in a file-scope device array instead of a caller
// workspace.  Each GPU holds its own copy of the array (at its own address) for as long as
// the process runs; host code reaches it only through the address hipGetSymbolAddress
// returns on that device, fetched once per device on the first call there.  Every call on
// a GPU shares that GPU's array: calls that overlap on one GPU (other streams, other
// threads) race.  bounds_partial writes the partials of every workgroup it launches before
// bounds_final reads them, so nothing needs clearing between calls.
namespace dsx {
constexpr int kScratchBlocks = 4096;        // partials held: 4096 x {min, max} = 64 KiB
constexpr int kScratchDevices = 64;
}  // namespace dsx

static __device__ __align__(256) double scratch[2 * dsx::kScratchBlocks];

namespace dsx {
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kScratchDevices];  // per device, set once under the mutex

static int scratch_address(double **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { record_hip_error(e); return DSX_ERR_HIP; }
    if (dev < 0 || dev >= kScratchDevices) return DSX_ERR_SHAPE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            record_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol);
            return DSX_ERR_HIP;
        }
        g_scratch_ptr[dev] = reinterpret_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return DSX_OK;
}
}  // namespace dsx

extern "C" int dsx_bounds_noalloc(const void *col, int32_t dtype, int64_t n, double *out2,
                                  void *stream) {
    if (!out2 || (n > 0 && !col)) return DSX_ERR_NULL;
    if (n < 0) return DSX_ERR_SHAPE;
    if (dtype != DSX_F32 && dtype != DSX_F64) return DSX_ERR_DTYPE;
    // dsx_bounds's launch shape, capped at the partials the array holds (the loops stride
    // over the grid, so any grid covers every row): no input needs more than the array
    int grid = grid_for(n < 1 ? 1 : n, kBoundsBlock, 8);
    grid = grid < kScratchBlocks ? grid : kScratchBlocks;
    double *part = nullptr;
    int st = scratch_address(&part);
    if (st) return st;
    hipStream_t s = as_stream(stream);
    if (dtype == DSX_F32)
        bounds_partial<float><<<grid, kBoundsBlock, 0, s>>>((const float *)col, n, part);
    else
        bounds_partial<double><<<grid, kBoundsBlock, 0, s>>>((const double *)col, n, part);
    st = launch_status();
    if (st) return st;
    bounds_final<<<1, kBoundsBlock, 0, s>>>(part, grid, out2);
    return launch_status();
}

namespace dsx {
int fill_async(void *plane, int32_t mode, int64_t ncells, hipStream_t s) {
    if (ncells < 0) return DSX_ERR_SHAPE;
    if (ncells == 0) return DSX_OK;
    if (!plane) return DSX_ERR_NULL;
    const int grid = grid_for(ncells, 256, 4);
    switch (mode) {
    case DSX_FILL_ZERO32: fill_kernel<uint32_t><<<grid, 256, 0, s>>>((uint32_t *)plane, 0u, ncells); break;
    case DSX_FILL_ONES32: fill_kernel<uint32_t><<<grid, 256, 0, s>>>((uint32_t *)plane, 0xFFFFFFFFu, ncells); break;
    case DSX_FILL_ZERO64: fill_kernel<uint64_t><<<grid, 256, 0, s>>>((uint64_t *)plane, 0ull, ncells); break;
    case DSX_FILL_ONES64: fill_kernel<uint64_t><<<grid, 256, 0, s>>>((uint64_t *)plane, ~0ull, ncells); break;
    default: return DSX_ERR_MODE;
    }
    return launch_status();
}
}  // namespace dsx

extern "C" int dsx_fill(void *plane, int32_t mode, int64_t ncells, void *stream) {
    return fill_async(plane, mode, ncells, as_stream(stream));
}

extern "C" int dsx_finalize(const dsx_finalize_desc *f, void *stream) {
    if (!f || !f->out) return DSX_ERR_NULL;
    if (f->ncells < 0) return DSX_ERR_SHAPE;
    if (f->mode < DSX_FIN_ANY || f->mode > DSX_FIN_SUMZERO) return DSX_ERR_MODE;
    if (f->mode == DSX_FIN_ANY) {
        if (!f->count) return DSX_ERR_NULL;
    } else if (f->mode == DSX_FIN_MIN || f->mode == DSX_FIN_MAX) {
        if (!f->key) return DSX_ERR_NULL;
        if (f->key_bytes != 4 && f->key_bytes != 8) return DSX_ERR_DTYPE;
    } else {
        if (!f->countv || !f->sum) return DSX_ERR_NULL;
    }
    if (f->ncells == 0) return DSX_OK;
    const int grid = grid_for(f->ncells, 256, 4);
    finalize_kernel<<<grid, 256, 0, as_stream(stream)>>>(*f);
    return launch_status();
}

// Partition combine (compiler.py:478-507 with the bases' _combine): counts and float64
// sums add, min / max keys take the unsigned min / max (empty cells hold the identity).
template <typename K>
__global__ __launch_bounds__(256) void combine_kernel(dsx_combine_desc c) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c.ncells; i += stride) {
        if (c.into.count) c.into.count[i] += c.part.count[i];
        if (c.into.countv) c.into.countv[i] += c.part.countv[i];
        if (c.into.sum) c.into.sum[i] += c.part.sum[i];
        if (c.into.min) {
            K *a = reinterpret_cast<K *>(c.into.min);
            const K b = reinterpret_cast<const K *>(c.part.min)[i];
            if (b < a[i]) a[i] = b;
        }
        if (c.into.max) {
            K *a = reinterpret_cast<K *>(c.into.max);
            const K b = reinterpret_cast<const K *>(c.part.max)[i];
            if (b > a[i]) a[i] = b;
        }
    }
}

extern "C" int dsx_combine(const dsx_combine_desc *c, void *stream) {
    if (!c) return DSX_ERR_NULL;
    if (c->ncells < 0) return DSX_ERR_SHAPE;
    if (c->key_bytes != 4 && c->key_bytes != 8) return DSX_ERR_DTYPE;
    if (c->into.mean || c->part.mean) return DSX_ERR_ACCUM;  // finalized means do not combine
    // the same planes must be present in `into` and `part`
    if ((c->into.count == nullptr) != (c->part.count == nullptr) ||
        (c->into.countv == nullptr) != (c->part.countv == nullptr) ||
        (c->into.sum == nullptr) != (c->part.sum == nullptr) ||
        (c->into.min == nullptr) != (c->part.min == nullptr) ||
        (c->into.max == nullptr) != (c->part.max == nullptr))
        return DSX_ERR_ACCUM;
    if (c->ncells == 0) return DSX_OK;
    const int grid = grid_for(c->ncells, 256, 4);
    if (c->key_bytes == 4) combine_kernel<uint32_t><<<grid, 256, 0, as_stream(stream)>>>(*c);
    else combine_kernel<uint64_t><<<grid, 256, 0, as_stream(stream)>>>(*c);
    return launch_status();
}

// where()'s lookup: the value of the selected row, NaN where no row was selected.
template <typename T>
__global__ __launch_bounds__(256) void take_kernel(dsx_take_desc t) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const T *col = reinterpret_cast<const T *>(t.col);
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < t.ncells; c += stride) {
        const int64_t i = t.index[c];
        t.out[c] = i < 0 ? __builtin_nan("") : (double)col[i];
    }
}

extern "C" int dsx_take(const dsx_take_desc *t, void *stream) {
    if (!t || !t->index || !t->out) return DSX_ERR_NULL;
    if (t->ncells < 0) return DSX_ERR_SHAPE;
    if (t->ncells == 0) return DSX_OK;
    if (!t->col) return DSX_ERR_NULL;
    const int grid = grid_for(t->ncells, 256, 4);
    hipStream_t s = as_stream(stream);
    switch (t->dtype) {
    case DSX_F32: take_kernel<float><<<grid, 256, 0, s>>>(*t); break;
    case DSX_F64: take_kernel<double><<<grid, 256, 0, s>>>(*t); break;
    case DSX_I8: take_kernel<int8_t><<<grid, 256, 0, s>>>(*t); break;
    case DSX_U8: take_kernel<uint8_t><<<grid, 256, 0, s>>>(*t); break;
    case DSX_I16: take_kernel<int16_t><<<grid, 256, 0, s>>>(*t); break;
    case DSX_I32: take_kernel<int32_t><<<grid, 256, 0, s>>>(*t); break;
    case DSX_I64: take_kernel<int64_t><<<grid, 256, 0, s>>>(*t); break;
    case DSX_U32: take_kernel<uint32_t><<<grid, 256, 0, s>>>(*t); break;
    default: return DSX_ERR_DTYPE;
    }
    return launch_status();
}

extern "C" int dsx_abi_version(void) { return DSX_ABI_VERSION; }

extern "C" const char *dsx_strerror(int status) {
    switch (status) {
    case DSX_OK: return "ok";
    case DSX_ERR_NULL: return "a required pointer is NULL";
    case DSX_ERR_DTYPE: return "unsupported dtype";
    case DSX_ERR_SHAPE: return "invalid shape, length or stride";
    case DSX_ERR_WORKSPACE: return "workspace missing or too small";
    case DSX_ERR_HIP: return "HIP runtime error";
    case DSX_ERR_ACCUM: return "accumulator planes inconsistent with the value column";
    case DSX_ERR_MODE: return "unknown mode";
    default: return "unknown status";
    }
}

extern "C" const char *dsx_last_hip_error(void) { return g_last_hip; }
