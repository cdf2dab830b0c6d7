/*
 * dsx.h -- C ABI of libdsx.so, the MI355X (gfx950) points / line aggregation library.
 *
 * Every pointer named "device" is a caller-owned device allocation (the Python host
 * passes torch-ROCm tensor data pointers); the library never allocates or frees caller
 * memory.  Every call is asynchronous on the given HIP stream (a hipStream_t passed as
 * void*; NULL = the default stream), returns DSX_OK (0) or a negative dsx_status, and
 * never throws.  Descriptors are validated on the host before anything is launched.
 * The library keeps no mutable global state other than the per-device address of
 * dsx_bounds_noalloc's array: calls on disjoint buffers are reentrant; calls sharing
 * planes or a workspace must be ordered (one stream).
 *
 * Correspondence with the reference (datashader 0.19.1) -- this ABI replaces the numba
 * kernels its backends call:
 *   dsx_points_agg     glyphs/points.py:188-221 (_perform_extend_points / extend_cpu /
 *                      extend_cuda) with the appends compiler.py:321-475 generates from
 *                      reductions.py (count 554-584, _sum_zero 958-962, sum 1056-1063,
 *                      min 1178-1182, max 1223-1227, any 842-861, by 789-798); with a
 *                      mean plane it is the m2 append of reductions.py:1128-1137
 *   dsx_lines_agg      glyphs/line.py:1033-1097 draw_segment (+ _liang_barsky 734-800,
 *                      map_onto_pixel_snap 688-720, _bresenham 986-1031) driven by the
 *                      extend loops of line.py:1099-1136 (axis 0), 1152-1193 (axis 0,
 *                      several columns), 1244-1290 (axis 1), 1340-1386 (axis 1, x
 *                      constant) and the y-constant variant of the same builder
 *   dsx_bounds         glyphs/glyph.py:66-78 (_compute_bounds_numba)
 *   dsx_finalize       reductions.py:491-493, 1092-1096, 1294-1297, 1312-1317, 1332-1337
 *                      (OptionalFieldReduction / sum / mean / var / std finalizers)
 *   dsx_fill           reductions.py:446-472 (the create functions' initial values)
 */
#ifndef DSX_H
#define DSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSX_ABI_VERSION 6

/* ---- status codes --------------------------------------------------------------- */
typedef enum {
    DSX_OK = 0,
    DSX_ERR_NULL = -1,        /* a required pointer is NULL */
    DSX_ERR_DTYPE = -2,       /* unsupported dtype code */
    DSX_ERR_SHAPE = -3,       /* non-positive canvas, negative length, bad stride */
    DSX_ERR_WORKSPACE = -4,   /* workspace missing or smaller than dsx_*_workspace_size */
    DSX_ERR_HIP = -5,         /* a HIP runtime call failed (launch or memory) */
    DSX_ERR_ACCUM = -6,       /* planes inconsistent with the value column */
    DSX_ERR_MODE = -7         /* unknown finalize / fill / categorizer mode */
} dsx_status;

/* ---- dtype codes ---------------------------------------------------------------- */
typedef enum {
    DSX_F32 = 1, DSX_F64 = 2,
    DSX_I8 = 3, DSX_U8 = 4, DSX_I16 = 5, DSX_I32 = 6, DSX_I64 = 7
} dsx_dtype;

/* ---- aggregate planes ----------------------------------------------------------- *
 * Each non-NULL plane is a dense device array of H*W*ncat cells, cell index
 * (y*W + x)*ncat + category (the reference's agg[y, x, cat] layout, C order):
 *   count   u32   rows in range                                  (count())
 *   countv  u32   rows in range whose value is not NaN           (count(v))
 *   sum     f64   sum of the non-NaN values (float64 accumulation; float32 values are
 *                 widened exactly)                                (_sum_zero(v))
 *   min/max u32 keys for a float32 value column, u64 keys for float64: order-preserving
 *           keys key(v) = bits(v) ^ (sign ? all-ones : sign-bit), so min/max are exact;
 *           an empty min cell holds DSX_MIN_EMPTY*, an empty max cell holds 0.
 * dsx_points_agg WRITES every cell of every requested plane (no initialisation needed).
 * dsx_lines_agg ADDS into planes the caller initialised with dsx_fill.               */
#define DSX_MIN_EMPTY32 0xFFFFFFFFu
#define DSX_MIN_EMPTY64 0xFFFFFFFFFFFFFFFFull

typedef struct {
    uint32_t *count;
    uint32_t *countv;
    double *sum;
    void *min;
    void *max;
    double *mean;     /* f64, already finalized: sum / count(v) of the non-NaN values, NaN
                         where a cell has none (mean's finalize, reductions.py:1294-1297);
                         dsx_points_agg / dsx_points_agg_indexed only                        */
} dsx_planes;

/* ---- view: the extend() arguments vt + bounds (points.py:223, line.py:362) --------- */
typedef struct {
    double sx, tx, sy, ty;        /* axis scale / translate (core.py:62-81)               */
    double xmin, xmax, ymin, ymax;/* inclusive data-space bounds                          */
    int64_t xxmax, yymax;         /* lines: round(m(xmax)*sx+tx), round(m(ymax)*sy+ty)    */
    int32_t width, height;        /* canvas W, H                                          */
    int32_t xlog, ylog;           /* 1: the axis maps v -> log10(v) before scaling        */
} dsx_view;                       /*    (LogAxis, core.py:127-145); bounds stay raw       */

/* ---- categorizer: by(...)'s category of a row (reductions.py:143-259) ------------- *
 * kind DSX_CAT_NONE    no categories (ncat must be 1)
 *      DSX_CAT_CODES   category = col[i] (pandas category codes; -1 wraps to ncat-1)
 *      DSX_CAT_MODULO  category = (col[i] - offset) mod modulo (Python sign rules)
 *      DSX_CAT_BINNING category = trunc((col[i] - bin0) / binsize), computed in float32
 *                      for a float32 column and float64 otherwise; < 0 -> bin_under,
 *                      >= nbins -> bin_over, NaN -> nbins
 * A category outside [0, ncat) after the wrap drops the row.                         */
typedef enum { DSX_CAT_NONE = 0, DSX_CAT_CODES = 1, DSX_CAT_MODULO = 2, DSX_CAT_BINNING = 3 }
    dsx_cat_kind;

typedef struct {
    int32_t kind;                 /* dsx_cat_kind                                         */
    int32_t dtype;                /* dtype of col                                         */
    const void *col;              /* device column (one element per row / line)           */
    int32_t ncat;                 /* categories (1 for DSX_CAT_NONE)                      */
    int32_t nbins, bin_under, bin_over;
    int64_t offset, modulo;
    double bin0, binsize;
} dsx_categorizer;

/* ---- points --------------------------------------------------------------------- *
 * How a call splits the rows between its passes (a performance choice; results do not
 * depend on it):
 *   WINDOW  a hot window of the canvas is aggregated in LDS during the scan; the other
 *           rows are spilled (cell + value) and aggregated per tile after a scatter by
 *           tile.  Best when a window holds most rows (clustered data).
 *   TILES   no window: the scan buckets every row by tile in LDS and writes full
 *           buckets as blocks of 16-bit tile-local offsets + values; the tile pass reads
 *           each tile's blocks.  Best when rows spread over the canvas (many categories,
 *           uniform data): one write and one read of 2 (+ value) bytes per row.
 *   AUTO    the call picks from the sampled rows: TILES when the best window would hold
 *           less than 70% of them and the canvas's tiles fit the scan's LDS buckets.
 * The strategy picked is kept in the window store (window[4]), so a caller can read it
 * back and pass it explicitly on later calls.  Another value: DSX_ERR_MODE.  A forced
 * TILES on a canvas whose tiles do not fit the buckets runs WINDOW.                    */
typedef enum { DSX_STRATEGY_AUTO = 0, DSX_STRATEGY_WINDOW = 1, DSX_STRATEGY_TILES = 2 }
    dsx_strategy;

typedef struct {
    int64_t n;                    /* rows                                                 */
    const void *x, *y;            /* device, n elements each                              */
    int32_t xy_dtype;             /* DSX_F32 | DSX_F64 (x and y share it)                 */
    int32_t val_dtype;            /* DSX_F32 | DSX_F64, 0 when val == NULL                */
    const void *val;              /* device value column or NULL                          */
    dsx_categorizer cat;
    /* Hot window (a performance hint; results do not depend on it): a device int32[8]
     * {x0, y0, w, h, strategy, 0, 0, 0} (pixels; the strategy a call ran), or NULL.
     * reuse_window = 0: the call picks the window (and, under AUTO, the strategy) from a
     * sample of the rows and, when window != NULL, stores it there; 1: the call uses the
     * stored record (an invalid window is treated as empty).  Either way every kernel of
     * the call reads the call's own copy in its workspace (copied in, or picked there and
     * copied out), so calls on several streams may share one store.                     */
    int32_t *window;
    int32_t reuse_window;
    int32_t strategy;             /* dsx_strategy                                         */
} dsx_points_desc;

/* Device workspace (bytes) dsx_points_agg needs for these inputs and planes. */
size_t dsx_points_workspace_size(const dsx_points_desc *d, const dsx_view *v,
                                 const dsx_planes *p);

/* Project every row to its bin and write the requested planes.  A row is dropped unless
 * xmin <= x <= xmax and ymin <= y <= ymax (NaN fails); its bin is trunc(x*sx + tx)
 * (float64, two separate roundings), clamped to W-1 at the upper edge.  Planes that read
 * the value column (countv, sum, min, max) require val.
 *
 * With mean != NULL (a device f64 plane of H*W*ncat cells) the call computes the m2 of
 * var / std instead: sum receives sum((v - mean[bin])^2) over the rows counted in
 * countv; only p->sum may be requested.  The reference's streaming update
 * (reductions.py:1128-1137) computes the same sum of squared deviations; this two-pass
 * form is its exact algebraic equivalent. */
int dsx_points_agg(const dsx_points_desc *d, const dsx_view *v, const dsx_planes *p,
                   const double *mean, void *workspace, size_t workspace_bytes, void *stream);

/* ---- lines ---------------------------------------------------------------------- *
 * nlines lines of nverts vertices each.  Vertex k of line l is
 *   x = X[l*x_line_stride + k*x_vert_stride],  y = Y[l*y_line_stride + k*y_vert_stride]
 * (element strides; a stride of 0 broadcasts), which expresses every layout of the
 * reference: axis 0 one line (1 line, vertex stride 1), axis 0 several columns, axis 1
 * (row-major or column-major), axis 1 with constant x or constant y.
 * Segment (l, k) -> (l, k+1) starts a new run when k == 0 (then segment_start =
 * plot_start for axis 0 and 1 for axis 1) or when vertex k-1 has a NaN coordinate.
 * The value / category index of a segment is k for axis 0 and l for axis 1.          */
#define DSX_PLOT_CONTEXT (-1)

typedef struct {
    int64_t nlines, nverts;
    const void *x, *y;
    int32_t x_dtype, y_dtype;     /* DSX_F32 | DSX_F64 each                               */
    int64_t x_line_stride, x_vert_stride, y_line_stride, y_vert_stride;
    int32_t axis;                 /* 0 or 1                                               */
    int32_t plot_start;           /* axis 0: 1 draw the first vertex, 0 do not;            */
                                  /* DSX_PLOT_CONTEXT: vertex 0 only precedes the line     */
                                  /* (segment 0 is not drawn, and vertex 0's NaN test      */
                                  /* gives segment 1 its segment_start) -- a chunk of a    */
                                  /* longer line uploaded behind its two previous rows     */
    const void *val; int32_t val_dtype;
    dsx_categorizer cat;          /* indexed like val                                     */
    const double *mean;           /* NULL, or per-cell means: then the only plane given,   */
                                  /* sum, accumulates (v - mean[cell])^2 of every pixel    */
                                  /* with a value (the var / std second pass)              */
} dsx_lines_desc;

int dsx_lines_agg(const dsx_lines_desc *d, const dsx_view *v, const dsx_planes *p,
                  void *stream);

/* ---- bounds (autorange) ----------------------------------------------------------- *
 * NaN-skipping min and max of n elements of a device column; writes {min, max} as two
 * doubles to the device pointer out2 (+inf/-inf when every element is NaN).  Needs a
 * device workspace of dsx_bounds_workspace_size(n) bytes. */
size_t dsx_bounds_workspace_size(int64_t n);
int dsx_bounds(const void *col, int32_t dtype, int64_t n, double *out2, void *workspace,
               size_t workspace_bytes, void *stream);

/* dsx_bounds without a workspace, for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The per-workgroup partials live in a 64 KiB array
 * inside the library (4096 partials; each GPU the library runs on holds its own copy
 * until the process ends); the call allocates nothing and n is unlimited, since the
 * launch never uses more workgroups than the array has partials.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (several streams or threads)
 * race and must be ordered by the caller.  Same results and status codes as dsx_bounds. */
int dsx_bounds_noalloc(const void *col, int32_t dtype, int64_t n, double *out2, void *stream);

/* ---- fill / finalize -------------------------------------------------------------- */
typedef enum {
    DSX_FILL_ZERO32 = 1,      /* u32 / f32 planes := 0                                    */
    DSX_FILL_ONES32 = 2,      /* u32 plane := 0xFFFFFFFF (min key, empty)                 */
    DSX_FILL_ZERO64 = 3,      /* u64 / f64 planes := 0                                    */
    DSX_FILL_ONES64 = 4       /* u64 plane := all ones (min key, empty)                   */
} dsx_fill_mode;

/* Fill ncells elements with a kernel (no hipMemset, so it is safe under graph capture). */
int dsx_fill(void *plane, int32_t mode, int64_t ncells, void *stream);

/* ---- row-index reductions (first / last / where) ----------------------------------- *
 * For every cell the row the reference's selector keeps (reductions.py:1340-1456,
 * 2236-2372, where 1842-2166): FIRST the smallest row index (_min_row_index; first()),
 * LAST the largest (_max_row_index; last()), ARGMAX / ARGMIN the smallest row index among
 * the rows whose value equals the cell's max / min (max._append / min._append keep the
 * first row that reaches the extreme).  Rows are positions 0..n-1 of the points desc;
 * rows with a NaN value are skipped when nan_check is set (always for ARGMAX / ARGMIN).
 * The categorizer of the desc splits cells by category as in dsx_points_agg.           */
typedef enum { DSX_RI_FIRST = 1, DSX_RI_LAST = 2, DSX_RI_ARGMAX = 3, DSX_RI_ARGMIN = 4 }
    dsx_ri_mode;

typedef struct {
    int32_t mode;                 /* dsx_ri_mode                                          */
    int32_t nan_check;            /* FIRST / LAST: skip rows whose value is NaN           */
    const void *key;              /* ARGMAX / ARGMIN: the max / min key plane of the value
                                     column as dsx_points_agg wrote it (key width = value
                                     width); NULL with a float32 value column and a
                                     workspace: one packed pass, no key plane needed      */
    int64_t *index;               /* out, H*W*ncat: the selected row, -1 where none       */
} dsx_rowindex_desc;

/* FIRST and LAST with a workspace of dsx_points_rowindex_workspace_size bytes take an LDS
 * pass: the rows in a hot window of the canvas (picked from a sample of the rows) keep
 * their cells' rows in LDS, the others take the global atomics; with workspace == NULL one
 * global-atomic pass.  ARGMAX / ARGMIN with a key plane take one streaming pass that keeps
 * the first row matching the key; without one (float32 values, a workspace) the LDS pass
 * over 64-bit cells key(v) << 32 | row (ARGMIN) or | ~row (ARGMAX), whose atomic min / max
 * keeps the extreme value's first row.  Same results every way. */
size_t dsx_points_rowindex_workspace_size(const dsx_points_desc *d, const dsx_view *v,
                                          const dsx_rowindex_desc *r);
int dsx_points_rowindex(const dsx_points_desc *d, const dsx_view *v,
                        const dsx_rowindex_desc *r, void *workspace, size_t workspace_bytes,
                        void *stream);

/* ---- n-value reductions (max_n / min_n / first_n / last_n) ------------------------ *
 * For every cell the rows of the reference's n-slot list (reductions.py:1457-1805,
 * 2375-2562), in its order: FIRST the first n rows (first_n, _min_n_row_index), LAST the
 * last n, latest first (last_n, _max_n_row_index), MAX / MIN the rows of the n largest /
 * smallest values, rows that tie keeping row order (max_n / min_n insert only before a
 * strictly smaller / larger value).  index[cell * n + k]; -1 pads cells with fewer rows.
 * Rows with a NaN value are skipped when nan_check is set (always for MAX / MIN).  n <= 8
 * (float32 values where values are read) takes n window passes, one per slot; otherwise
 * one radix sort of the kept rows.  The workspace size depends on which.               */
typedef enum { DSX_NS_FIRST = 1, DSX_NS_LAST = 2, DSX_NS_MAX = 3, DSX_NS_MIN = 4 }
    dsx_ns_mode;

typedef struct {
    int32_t mode;                 /* dsx_ns_mode                                          */
    int32_t n;                    /* rows kept per cell, 1..4096                          */
    int32_t nan_check;
    int32_t reserved;
    int64_t *index;               /* out, H*W*ncat*n                                      */
} dsx_nselect_desc;

size_t dsx_points_nselect_workspace_size(const dsx_points_desc *d, const dsx_view *v,
                                         const dsx_nselect_desc *r);
int dsx_points_nselect(const dsx_points_desc *d, const dsx_view *v, const dsx_nselect_desc *r,
                       void *workspace, size_t workspace_bytes, void *stream);

/* out[c] = col[index[c]] as float64, NaN where index[c] < 0 (where()'s lookup,
 * reductions.py:1914-1925).  col dtype: DSX_F32, DSX_F64, DSX_I8..DSX_I64, DSX_U32.     */
typedef struct {
    int64_t ncells;
    const int64_t *index;
    const void *col;
    int32_t dtype;
    int32_t reserved;
    double *out;
} dsx_take_desc;

int dsx_take(const dsx_take_desc *t, void *stream);

/* ---- partition combine ---------------------------------------------------------------- *
 * into[p] (+)= part[p] for every plane p present in both (the planes present must be the
 * same): count / countv / sum add, min / max keys take the unsigned min / max.  Replaces
 * the combine step of compiler.py:478-507 (count._combine, _sum_zero._combine, min/max
 * _combine, reductions.py:652-654, 985-996, 1197-1205, 1252-1260) that dask uses to merge
 * partition aggregates (data_libraries/dask.py:168-217); here it merges the planes of the
 * chunks of a host-resident table streamed through the GPU.                              */
typedef struct {
    int64_t ncells;
    int32_t key_bytes;            /* width of the min / max keys: 4 or 8                  */
    int32_t reserved;
    dsx_planes into;              /* accumulated in place                                 */
    dsx_planes part;              /* one chunk's planes                                   */
} dsx_combine_desc;

int dsx_combine(const dsx_combine_desc *c, void *stream);

typedef enum {
    DSX_FIN_ANY = 1,          /* out u8  = count > 0                                      */
    DSX_FIN_SUM = 2,          /* out f64 = countv > 0 ? sum : NaN                         */
    DSX_FIN_MEAN = 3,         /* out f64 = countv > 0 ? sum / countv : NaN                */
    DSX_FIN_VAR = 4,          /* out f64 = countv > 0 ? m2 / countv : NaN   (m2 in sum)   */
    DSX_FIN_STD = 5,          /* out f64 = countv > 0 ? sqrt(m2 / countv) : NaN           */
    DSX_FIN_MIN = 6,          /* out f64 = decoded min key, NaN when empty                */
    DSX_FIN_MAX = 7,          /* out f64 = decoded max key, NaN when empty                */
    DSX_FIN_SUMZERO = 8       /* out f64 = sum (0 when empty)                             */
} dsx_finalize_mode;

typedef struct {
    int32_t mode;             /* dsx_finalize_mode                                        */
    int32_t key_bytes;        /* DSX_FIN_MIN / MAX: 4 (float32 values) or 8 (float64)     */
    int64_t ncells;
    const uint32_t *count;    /* DSX_FIN_ANY: the count (or countv) plane                 */
    const uint32_t *countv;
    const double *sum;        /* sum, or m2 for DSX_FIN_VAR / DSX_FIN_STD                 */
    const void *key;          /* min / max key plane                                      */
    void *out;                /* device output (u8 or f64)                                */
} dsx_finalize_desc;

int dsx_finalize(const dsx_finalize_desc *f, void *stream);

/* ---- serving: grid index over a resident point set --------------------------------- *
 * No counterpart in the reference: the reference scans every row on every call.  A
 * viewport server keeps its point set resident, sorted by the cell of a grid x grid
 * lattice over the set's extent, so that one viewport reads only the rows of the cells
 * it overlaps.  dsx_grid_index_build sorts the rows (x, y and an optional value column,
 * all of one dtype) by cell -- rows with a NaN coordinate go last, outside every cell --
 * writes the sorted columns and cell_start[grid*grid + 1]: the rows of cell
 * (gy*grid + gx) are [cell_start[c], cell_start[c+1]).  A row's cell is
 * min(grid-1, (int)((x - x0) * (grid / (x1 - x0)))) (the same for y), so a viewport
 * must take one extra cell on each side of the cells its range maps to. */
typedef struct {
    int64_t n;
    int32_t grid;             /* cells per side, 1 .. 4096                                */
    int32_t dtype;            /* DSX_F32 or DSX_F64 for x, y and val                       */
    const void *x, *y, *val;  /* device columns (val may be NULL)                          */
    double x0, x1, y0, y1;    /* the lattice's extent (x1 > x0, y1 > y0)                   */
    void *out_x, *out_y, *out_val;  /* device, n rows each (out_val NULL without val)      */
    int64_t *cell_start;      /* device [grid*grid + 1]                                   */
} dsx_grid_index_desc;

size_t dsx_grid_index_workspace_size(const dsx_grid_index_desc *g);
int dsx_grid_index_build(const dsx_grid_index_desc *g, void *workspace,
                         size_t workspace_bytes, void *stream);

/* The lattice of a built index, as dsx_points_agg_indexed reads it. */
typedef struct {
    int32_t grid;             /* cells per side, 1 .. 1024                                */
    int32_t reserved;
    double x0, x1, y0, y1;    /* the lattice's extent, as given to dsx_grid_index_build    */
    const int64_t *cell_start;  /* device [grid*grid + 1]                                 */
} dsx_grid_ref;

/* dsx_points_agg over the sorted columns of an index (d->x, d->y, d->val are the
 * out_x / out_y / out_val of dsx_grid_index_build, d->n its row count): the canvas is
 * cut into tiles whose planes fit LDS and each tile reads only the rows of the lattice
 * cells under it, so no row is spilled or partitioned.  Same planes, results and error
 * codes as dsx_points_agg; linear axes (xlog = ylog = 0) and no categorizer. */
size_t dsx_points_indexed_workspace_size(const dsx_points_desc *d, const dsx_view *v,
                                         const dsx_planes *p, const dsx_grid_ref *g);
int dsx_points_agg_indexed(const dsx_points_desc *d, const dsx_view *v, const dsx_planes *p,
                           const double *mean, const dsx_grid_ref *g, void *workspace,
                           size_t workspace_bytes, void *stream);

/* Copy row ranges of up to four columns into contiguous destinations: item k of the
 * device array items[3*nitems] is (src_row, dst_row, rows). */
typedef struct {
    int32_t ncols;
    int32_t elem_bytes[4];    /* 4 or 8                                                   */
    const void *src[4];
    void *dst[4];
    const int64_t *items;
    int64_t nitems;
} dsx_gather_desc;

int dsx_gather_rows(const dsx_gather_desc *g, void *stream);

/* ---- batched viewports (the served form) ------------------------------------------- *
 * No counterpart in the reference, which scans every row on every call
 * (core.py:1334-1359 -> data_libraries/pandas.py:23-66).  An engine holds one resident
 * set sorted by dsx_grid_index_build (float32 x, y, value) and answers batches of
 * viewport requests -- each its own ranges (linear axes), canvas size and output planes
 * -- with one replay of a HIP graph captured at dsx_batch_create: a fixed kernel
 * sequence driven by a work queue of (request, canvas tile, row part) items that
 * persistent workgroups drain.  Each request's planes are those of dsx_points_agg over
 * the whole set for count() / count(v) / sum(v) / mean(v) (reductions.py 582-584,
 * 554-558, 1056-1063 with 1092-1096, 1294-1297), finalized: count and count(v) u32, sum
 * f64 (NaN where no value), mean f64 (NaN where no value).  Every output cell is
 * written, so outputs need no initialisation.                                          */
#define DSX_BATCH_COUNT 1
#define DSX_BATCH_COUNTV 2
#define DSX_BATCH_SUM 4
#define DSX_BATCH_MEAN 8

typedef struct {
    double sx, tx, sy, ty;        /* the request's view (dsx_view, linear axes, sx, sy > 0) */
    double xmin, xmax, ymin, ymax;
    int32_t width, height;
    int32_t tile_w, tile_h;       /* in: tile shape in pixels (0: default); dsx_batch_submit */
                                  /* narrows it to the LDS tile                             */
    int64_t part_rows;            /* rows per work item of a heavy tile (0: 65536)          */
    int32_t tile_base;            /* set by dsx_batch_submit                                */
    int32_t reserved;
    uint32_t *count, *countv;     /* device outputs, H*W each (those the engine produces)   */
    double *sum, *mean;
} dsx_batch_req;

typedef struct {
    int64_t n;                    /* resident rows, < 2^32                                  */
    const float *x, *y, *val;     /* the index's sorted columns (val NULL: count only)      */
    int32_t grid;                 /* the index's lattice, 1 .. 4096                         */
    uint32_t planes;              /* DSX_BATCH_* bits                                       */
    double x0, x1, y0, y1;        /* the lattice's extent                                   */
    const int64_t *cell_start;    /* device [grid*grid + 1]                                 */
    int32_t max_requests;         /* requests per batch                                     */
    int32_t max_tiles;            /* canvas tiles per batch (all requests)                  */
    int32_t max_regions;          /* scratch regions for the parts of split tiles           */
    int32_t threads;              /* aggregation workgroup: 256, 512 or 1024 threads        */
    int32_t workgroups;           /* persistent aggregation workgroups                      */
    int32_t reserved;
} dsx_batch_config;

typedef struct dsx_batch dsx_batch;

/* Device workspace of an engine, the pinned host table the caller writes requests into
 * (dsx_batch_table_bytes of page-locked memory, e.g. from hipHostMalloc: every batch
 * copies it to the workspace asynchronously), and the cells of one LDS tile. */
size_t dsx_batch_workspace_size(const dsx_batch_config *c);
size_t dsx_batch_table_bytes(int32_t max_requests);
int32_t dsx_batch_tile_cells(const dsx_batch_config *c);
/* Create an engine over caller-owned workspace and pinned table; with use_graph the
 * kernel sequence is captured once into a HIP graph (replayed by every submit), else it
 * is launched kernel by kernel. */
int dsx_batch_create(const dsx_batch_config *c, void *workspace, size_t workspace_bytes,
                     void *host_table, int32_t use_graph, dsx_batch **out);
/* The request slots of the pinned table (max_requests of them). */
dsx_batch_req *dsx_batch_table(dsx_batch *b);
/* Validate the first nreq requests written into the table and run the batch on stream.
 * The table must not be rewritten until the batch has completed. */
int dsx_batch_submit(dsx_batch *b, int32_t nreq, void *stream);
int dsx_batch_destroy(dsx_batch *b);

/* ---- shade: aggregate planes -> RGBA images ---------------------------------------- *
 * Replaces transfer_functions/__init__.py shade (616-747) for device-resident planes:
 *   DSX_SHADE_CMAP       _interpolate (251-359) with a list colormap: colour channels
 *                        interpolated over linspace(span, ncolors), alpha constant
 *   DSX_SHADE_COLOR      _interpolate with one colour: alpha interpolated over
 *                        linspace(span, nalpha) onto arange(min_alpha, alpha + 1)
 *   DSX_SHADE_CATEGORICAL _colorize (360-463) + _interpolate_alpha (466-538): an
 *                        [H, W, ncat] plane, colours mixed by value, alpha from the total
 *   DSX_SHADE_COLOR_KEY  _apply_discrete_colorkey (541-613): a 2D plane of category values
 * how: eq_hist (148-223), log1p, cube root or linear (225-228).  `out` receives H*W pixels
 * r | g << 8 | b << 16 | a << 24 (the reference's uint32 view of RGBA bytes).
 * Every scalar the reference derives from the data (offset, span, histogram, cdf) is
 * computed on the device: a call launches a fixed kernel sequence with no host
 * synchronisation.  Data dtypes: DSX_U32, DSX_BOOL (u8 0/1), DSX_I32, DSX_I64, DSX_F32,
 * DSX_F64 (integers exact up to 2^53).                                                  */
#define DSX_U32 8
#define DSX_BOOL 9

typedef enum {
    DSX_SHADE_CMAP = 0, DSX_SHADE_COLOR = 1, DSX_SHADE_CATEGORICAL = 2, DSX_SHADE_COLOR_KEY = 3
} dsx_shade_kind;

typedef enum {
    DSX_HOW_LINEAR = 0, DSX_HOW_LOG = 1, DSX_HOW_CBRT = 2, DSX_HOW_EQ_HIST = 3
} dsx_how;

typedef struct {
    int32_t kind;            /* dsx_shade_kind */
    int32_t dtype;           /* dtype code of `data` */
    int64_t H, W, ncat;      /* ncat: categories (CATEGORICAL), else 1 */
    const void *data;        /* device, H*W*ncat, C order */
    uint32_t *out;           /* device, H*W pixels */
    int32_t how;             /* dsx_how */
    int32_t ncolors;         /* rows of `colors` (and of `keys` for COLOR_KEY) */
    const uint8_t *colors;   /* device, ncolors x (r, g, b) */
    const double *keys;      /* device, COLOR_KEY: the category value of each colour */
    double alpha, min_alpha; /* as given */
    int32_t nalpha;          /* COLOR: len(arange(min_alpha, alpha + 1)) */
    int32_t has_span;        /* span given: clip + fixed normalisation */
    double span_lo, span_hi; /* span as given */
    double span_delta;       /* span[1] - span[0] as the caller evaluated it */
    int32_t rescale_discrete_levels;
    int32_t has_baseline;    /* CATEGORICAL: color_baseline given; COLOR_KEY: likewise */
    double baseline;
} dsx_shade_desc;

size_t dsx_shade_workspace_size(const dsx_shade_desc *d);
int dsx_shade(const dsx_shade_desc *d, void *workspace, size_t workspace_bytes, void *stream);

/* ---- misc --------------------------------------------------------------------------- */
int dsx_abi_version(void);
const char *dsx_strerror(int status);
/* Last HIP error string recorded by a call that returned DSX_ERR_HIP on this thread. */
const char *dsx_last_hip_error(void);

#ifdef __cplusplus
}
#endif
#endif /* DSX_H */
