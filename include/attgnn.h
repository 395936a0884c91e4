/*
 * attgnn.h -- C ABI of the MI355X window-block sparse GNN library (libattgnn.so).
 *
 * Plain pointers and sizes only.  Device pointers are HIP device memory on the
 * current device; every GPU entry point takes the hipStream_t it runs on
 * (pass the caller's current stream) and returns 0 or a nonzero status code
 * (an ATTGNN_E_* value, or a hipError_t for launch failures).  Outputs are
 * allocated by the caller.  No entry point allocates device memory, keeps
 * global state, or synchronises the stream.  Results are deterministic: a
 * window split across workgroups is reduced in a fixed order.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the reference repository root).
 */
#ifndef ATTGNN_H
#define ATTGNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *attgnn_stream_t; /* == hipStream_t */

/* ---- status codes -------------------------------------------------------- */
enum {
    ATTGNN_OK = 0,
    ATTGNN_E_ARG = -1,      /* bad shape, size, pointer or enum value          */
    ATTGNN_E_FORMAT = -2,   /* malformed CSR / block structure                 */
    ATTGNN_E_NOMEM = -3,    /* host allocation failed (preprocessing)          */
    ATTGNN_E_UNSUPPORTED = -4,
    ATTGNN_E_SCRATCH = -5,  /* larger than the fixed scratch of an allocation-free
                               entry point                                       */
};
const char *attgnn_status_string(int status);

/* ---- element types -------------------------------------------------------- */
enum {
    ATTGNN_F32 = 0,
    ATTGNN_F16 = 1,
    ATTGNN_BF16 = 2,
    ATTGNN_FP8E4M3 = 3, /* OCP e4m3fn */
    ATTGNN_I32 = 4,
};

/* ---- multi-head dense layouts ------------------------------------------- */
enum {
    ATTGNN_LAYOUT_HMD = 0, /* [heads][rows][dim]  ("h_m" in the reference)   */
    ATTGNN_LAYOUT_MHD = 1, /* [rows][heads][dim]  ("m_h" in the reference)   */
};

/* ---- SpMM arithmetic modes ----------------------------------------------- */
enum {
    ATTGNN_MODE_NATIVE = 0, /* MFMA in the tile/X element type, FP32 accumulate:
                               f16 x f16, bf16 x bf16, fp8 x fp8 (+ scales)     */
    ATTGNN_MODE_BF16X3 = 1, /* fp32 tile/X split into bf16 hi+lo, 3 MFMA terms  */
    ATTGNN_MODE_F32 = 2,    /* fp32 MFMA (16x16x4)                              */
};

/* ====================================================================== */
/* Window-block preprocessing (host, OpenMP)                               */
/* ====================================================================== */

/* Source of the tile contents. */
enum {
    ATTGNN_SRC_DEG_F16 = 0, /* value = Half(float(d[row]) * float(d[col])), d fp16
                               (gcn/Block_gcn/example.cpp:105)                  */
    ATTGNN_SRC_DEG_F32 = 1, /* value = cvt(d[row] * d[col]) in fp32, d fp32
                               (example.cpp:259 when out type is F32)           */
    ATTGNN_SRC_EDGE_F32 = 2,/* value = cvt(v[m]) for CSR position m             */
    ATTGNN_SRC_MAP = 3,     /* value = CSR position m (int32), -1 = padding     */
};

typedef struct attgnn_blocks attgnn_blocks; /* opaque host-side result */

/*
 * Build the window-block structure of a CSR matrix.
 *
 * Replaces Block_sr_bcrs.blockProcess_fp16_balance  (gcn/Block_gcn/example.cpp:48)
 *          Block_sr_bcrs.blockProcess_tf32_balance  (gcn/Block_gcn/example.cpp:203)
 *          ATT_Block[_v3].blockProcess_sddmm_balance_gnn[_trans]
 *                                                   (Block/include/example.h:4-6)
 *
 * row[rows_plus1], col[nnz]: CSR (columns need not be sorted).  src: one of
 * ATTGNN_SRC_*; src_ptr: d[rows] (fp16 bits or fp32) or v[nnz] (fp32) or NULL.
 * out_type: tile element type (F16/BF16/F32 for values, I32 for MAP).
 * window rows per window, wide columns per block, part_size blocks per part.
 * emit_empty: 1 = an empty window emits the part (0,0) (tf32 variant),
 *             0 = it emits nothing (fp16 / sddmm variants).
 * want_trans: MAP only -- also build the transposed-edge map.
 * nthreads: OpenMP threads (0 = default).
 */
int attgnn_blocks_build(const int32_t *row, int64_t rows_plus1, const int32_t *col, int64_t nnz,
                        int src, const void *src_ptr, int out_type, int window, int wide,
                        int part_size, int emit_empty, int want_trans, int nthreads,
                        attgnn_blocks **out);

/* Sizes of the arrays: rp[n_rp], col[n_col], val[n_val], map_t[n_val] (if
 * built), window_of_part[n_parts], atomic[n_parts]; U = sum over windows of
 * distinct columns (unpadded). */
int attgnn_blocks_sizes(const attgnn_blocks *b, int64_t *n_rp, int64_t *n_col, int64_t *n_val,
                        int64_t *n_parts, int64_t *n_unique, int *has_map_t);

/* Copy the arrays out into caller buffers (any of them may be NULL). */
int attgnn_blocks_export(const attgnn_blocks *b, int32_t *rp, int32_t *col, void *val,
                         int32_t *map_t, int32_t *window_of_part, int32_t *atomic);

void attgnn_blocks_free(attgnn_blocks *b);

/*
 * Compact tiles of a structure built with wide = 32 and window <= 16:
 * masks[C] (uint16, bit r = window row r has an entry in that condensed
 * column), block_base[C/32 + 1] (int64 offsets of each 32-column block's
 * values), values[n_values] (tile type, or int32 edge ids for a map) packed
 * per block, per 8-column group, row-major within a group, and optionally
 * positions[n_values] (uint16 r*32 + c of each entry in its block).  An entry
 * is a tile value whose bits are not all zero, or a map id >= 0.
 *
 * No counterpart in the reference, which stores dense window x wide tiles
 * (gcn/Block_gcn/example.cpp:84-106); the reference format stays available.
 */
int attgnn_blocks_compact_sizes(const attgnn_blocks *b, int64_t *n_blocks, int64_t *n_values);
int attgnn_blocks_compact_export(const attgnn_blocks *b, uint16_t *masks, int64_t *block_base, void *values,
                                 uint16_t *positions);

/* ====================================================================== */
/* Window-block preprocessing (device)                                     */
/* ====================================================================== */

/*
 * The same structure as attgnn_blocks_build, built on the GPU from a CSR that
 * is already in device memory; every output is bit-identical to the host
 * build (same sources, element types, window / wide / part_size, empty-window
 * rule, unsorted rows and duplicate entries).
 *
 * Replaces ATT_Block_gpu.preprocess_gpu_fs  (Block_gpu/setup.py:10-14; caller
 *            end2end_overhead/gat_final/mgat_csr_ov/mdataset_fp16.py:53-58),
 *          and on the device the host windowing of
 *            Block_sr_bcrs.blockProcess_{fp16,tf32}_balance (gcn/Block_gcn/example.cpp:48,203)
 *            ATT_Block[_v3].blockProcess_sddmm_balance_gnn[_trans] (Block/include/example.h:4-6)
 *
 * Two phases, both asynchronous on `stream`:
 *  plan: checks the CSR and finds every window's distinct columns, writing
 *    into totals (device int64[8]): [0] status (ATTGNN_OK, ATTGNN_E_FORMAT for a
 *    malformed CSR, ATTGNN_E_UNSUPPORTED if the column count exceeds int32),
 *    [1] n_rp, [2] n_col, [3] n_val, [4] n_parts, [5] n_unique (as
 *    attgnn_blocks_sizes).  The caller reads totals, allocates the outputs
 *    and calls
 *  fill: with the same row, col, window, wide and part_size and the same
 *    workspace (unchanged since the plan); map_t (MAP only, may be NULL)
 *    receives the transposed-edge map.  col_out and the part table (rp, with
 *    window_of_part and atomic) may be NULL, and are then not written: a
 *    second fill of the same plan in another tile type needs only the tiles
 *    (window_of_part / atomic require rp).
 * workspace: device memory of attgnn_blocks_device_workspace_bytes() bytes.
 * row[rows_plus1], col[nnz], src_ptr: device pointers.
 */
int64_t attgnn_blocks_device_workspace_bytes(int64_t rows_plus1, int64_t nnz, int window);

int attgnn_blocks_plan_device(const int32_t *row, int64_t rows_plus1, const int32_t *col, int64_t nnz,
                              int window, int wide, int part_size, int emit_empty, void *workspace,
                              int64_t workspace_bytes, int64_t *totals, attgnn_stream_t stream);

int attgnn_blocks_fill_device(const int32_t *row, int64_t rows_plus1, const int32_t *col, int64_t nnz,
                              int src, const void *src_ptr, int out_type, int window, int wide,
                              int part_size, const void *workspace, int64_t workspace_bytes,
                              int32_t *rp, int32_t *col_out, void *val, int32_t *map_t,
                              int32_t *window_of_part, int32_t *atomic, attgnn_stream_t stream);

/*
 * perm[map[s]] = map_t[s] for every tile slot s with map[s] >= 0: from the
 * edge-id map and transposed map of a structure that covers every CSR entry
 * once (rows padded to the window, no duplicate entries), the CSR position of
 * the transpose of each entry (-1 where the transposed entry is absent);
 * entries no slot covers are left untouched.
 *
 * Replaces the column-sum indexing of ATT_SpMM_v{1,3}.spmm_*_h_m_rowsum called
 * with degrees_trans (end2end_head/gat_final/mgat_csr/mgat_conv.py:42-48).
 */
int attgnn_map_transpose_perm_device(const int32_t *map, const int32_t *map_t, int64_t n_val, int64_t nnz,
                                     int64_t *perm, attgnn_stream_t stream);

/*
 * CSR of A^T on the device: t_rowptr[n_cols+1], t_col[nnz] (the row of A of
 * each entry, ascending within a row of A^T) and t_perm[nnz] (the CSR position
 * in A of each entry), equal to a stable host sort of A's entries by (column,
 * row).  status (device int64[1]) receives ATTGNN_OK or ATTGNN_E_FORMAT (row
 * pointers not a partition of [0, nnz), or a column outside [0, n_cols)).
 * workspace: attgnn_csr_transpose_workspace_bytes() bytes.
 *
 * No counterpart in the reference: its layers reuse A for A^T
 * (end2end_ori/gcn_final/mgat_csr/mgat_conv.py:57-74), exact only for a
 * symmetric A.
 */
int64_t attgnn_csr_transpose_workspace_bytes(int64_t n_rows, int64_t n_cols, int64_t nnz);
int attgnn_csr_transpose_device(const int32_t *rowptr, const int32_t *col, int64_t n_rows, int64_t n_cols,
                                int64_t nnz, int32_t *t_rowptr, int32_t *t_col, int64_t *t_perm, int64_t *status,
                                void *workspace, int64_t workspace_bytes, attgnn_stream_t stream);

/* out[s] = map[s] < 0 ? -1 : (int32) remap[map[s]]: edge ids of a transposed
 * structure's map re-expressed in A's edge order (remap = t_perm). */
int attgnn_map_remap_device(const int32_t *map, int64_t n, const int64_t *remap, int64_t remap_n, int32_t *out,
                            attgnn_stream_t stream);

/*
 * Compact tiles on the device, bit-identical to attgnn_blocks_compact_export:
 * count writes masks[n_cols] and block_base[n_cols/32 + 1] (block_base[n_cols/32]
 * = n_values, which the caller reads to size values / positions); fill packs
 * values[n_values] and (if not NULL) positions[n_values].  val holds the
 * n_cols * window tiles of a structure built with wide = 32 (tile_type
 * F16 / BF16 / F32, or I32 for an edge-id map).  count takes a device
 * workspace of attgnn_blocks_compact_workspace_bytes(n_cols) bytes (the block
 * counts and the device-wide scan's storage).  val and masks must be 16-byte
 * aligned (tiles and masks are read with 16-byte loads); ATTGNN_E_ARG if not.
 *
 * No counterpart in the reference (see attgnn_blocks_compact_sizes).
 */
int64_t attgnn_blocks_compact_workspace_bytes(int64_t n_cols);
int attgnn_blocks_compact_count_device(const void *val, int tile_type, int64_t n_cols, int window,
                                       uint16_t *masks, int64_t *block_base, void *workspace,
                                       int64_t workspace_bytes, attgnn_stream_t stream);
int attgnn_blocks_compact_fill_device(const void *val, int tile_type, int64_t n_cols, int window,
                                      const uint16_t *masks, const int64_t *block_base, void *values,
                                      uint16_t *positions, attgnn_stream_t stream);

/*
 * Canonical SpMM task table of a window-block structure (device).  Each
 * window's condensed columns, from its start to its padded end, are cut into
 * chunks of chunk_blocks * wide columns (at least one task per window, empty
 * windows included); the tasks are written in the structure's own part format
 * -- task_rp [2*max_tasks+1] running (real, pad) pairs, task_window[max_tasks],
 * task_atomic[max_tasks] = 1 when the window has more than one chunk -- in
 * window order, slots past the total with window -1 and empty column ranges
 * (which every attgnn_spmm_* entry point skips).  *n_tasks (device int64)
 * receives the total.  Run over these tasks instead of the parts, a window's
 * SpMM sum is grouped by the chunk size alone: results are bitwise the same
 * whatever partSize split the structure.  max_tasks >= attgnn_blocks_tasks_max
 * (parts + ceil(n_cols / (chunk_blocks*wide))); workspace of
 * attgnn_blocks_tasks_workspace_bytes(parts) bytes.  Parts with a negative
 * window are dropped.
 *
 * No counterpart in the reference (whose sums are grouped by partSize,
 * gcn/Block_gcn/example.cpp:120-151, and accumulated with atomicAdd,
 * gcn/SpMM/src/spmmKernel.cu:226-262).
 */
int64_t attgnn_blocks_tasks_max(int64_t parts, int64_t n_cols, int wide, int chunk_blocks);
int64_t attgnn_blocks_tasks_workspace_bytes(int64_t parts);
int attgnn_blocks_tasks_device(const int32_t *rp, const int32_t *window_of_part, int64_t parts, int64_t n_cols,
                               int wide, int chunk_blocks, int32_t *task_rp, int32_t *task_window,
                               int32_t *task_atomic, int64_t max_tasks, void *workspace, int64_t workspace_bytes,
                               int64_t *n_tasks, attgnn_stream_t stream);

/* ====================================================================== */
/* Window-block SpMM (device)                                              */
/* ====================================================================== */

/*
 * Y[m_ori, dim_n] (row-major, y_type F32/F16/BF16) = A . X, A given as window blocks;
 * accumulation is FP32 in every mode, Y is rounded once at the store.
 *
 * Replaces SpMM_sr_bcrs.forward_fp16_gnn  (gcn/SpMM/src/benchmark.cpp:35-44,
 *                                          launcher spmmKernel.cu:1693-1727)
 *          SpMM_sr_bcrs.forward_tf32_gnn  (benchmark.cpp:91-100,
 *                                          launcher spmmKernel.cu:1731-1765)
 *
 * rp[2*parts+1], col[...], val[...] (tile_type), window_of_part[parts],
 * atomic[parts] exactly as produced by attgnn_blocks_build.  X[n_x, dim_n] of
 * x_type, row-major.  window: 8 or 16.  wide: 4, 8, 16 or 32.
 * row_scale[m_ori] / col_scale[dim_n]: FP8 mode scales (NULL otherwise).
 * workspace: fp32 scratch of at least attgnn_spmm_workspace_elems(...) elements
 * (partial sums of split windows, reduced in part order); required when
 * parts > 0.  Every row of Y is written (rows no part covers are zero).
 * X must be aligned to its lane vector (<= 16 bytes) and Y to 16 bytes.
 * A part whose window_of_part is negative is padding: it does no work and
 * stores nothing (atomic must be 0 there).  Buffers sized for a capacity and
 * filled per batch -- a captured graph replayed over batches of different
 * sizes -- mark their unused parts so.  The same holds for every
 * attgnn_spmm_* entry point below.  No counterpart in the reference.
 */
int attgnn_spmm_block(const int32_t *rp, const int32_t *col, const void *val, int tile_type,
                      const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                      const void *X, int x_type, int64_t n_x, void *Y, int y_type, int64_t m_ori,
                      int64_t dim_n, int window, int wide, int mode, const float *row_scale,
                      const float *col_scale, float *workspace, int64_t workspace_elems,
                      int flags, attgnn_stream_t stream);

/* attgnn_spmm_block flags */
enum {
    ATTGNN_SPMM_ROWS_COVERED = 1, /* every row < m_ori lies in a window that has a
                                     part (built with emit_empty=1 and m_ori <=
                                     windows*window): Y is not pre-zeroed       */
    /* k-loop schedule (bits 4-5); results are bitwise identical across them */
    ATTGNN_SPMM_PIPE_DEFAULT = 0 << 4, /* PAIR for 8/16-bit dense or compact
                                          tiles, and for 8/16-bit edge values of
                                          an edge-id map when the launch has
                                          fewer than 16,384 (part, head, slice)
                                          tasks; else ONE                        */
    ATTGNN_SPMM_PIPE_ONE = 1 << 4,     /* next block's column ids (and compact
                                          masks) prefetched                      */
    ATTGNN_SPMM_PIPE_PAIR = 2 << 4,    /* two blocks per iteration + prefetch    */
    ATTGNN_SPMM_PIPE_NONE = 3 << 4,    /* no prefetch                            */
    ATTGNN_SPMM_NO_XCD_SLICES = 1 << 6, /* by default, when the feature-slice count
                                          divides 8, every XCD works on a single
                                          slice so its L2 caches one slice of the
                                          gathered rows; this flag turns it off  */
    /* bits 8-11: features per lane (1, 2, 4 or 8; 0 = automatic), i.e. slices of
       16x that many features; only narrower than the automatic choice        */
    /* bits 12-14 (attgnn_spmm_compact_striped): log2(owner groups) + 1, 0 =
       automatic (the fewest leaving each XCD about 16 MB of X); at most
       min(owners, 8 / slices)                                                  */
};

int64_t attgnn_spmm_workspace_elems(int64_t parts, int window, int64_t dim_n);

/*
 * attgnn_spmm_block without the workspace argument: the same arguments
 * otherwise and the same results bit for bit, and it allocates nothing (no
 * hipMalloc / hipFree, no allocating library routine).  The scratch is a fixed
 * device array of ATTGNN_NOALLOC_SCRATCH_ELEMS floats (64 MiB) that every GPU
 * the library runs on holds until the process ends; a problem needing more
 * (attgnn_spmm_workspace_elems(parts, window, dim_n) elements, i.e. about
 * parts * window * dim_n) returns ATTGNN_E_SCRATCH -- for window 16, fewer
 * than 4,096 parts at 256 features or 8,192 at 128.  All calls on one GPU share the
 * array: calls that overlap in time on one GPU race (serialise them, e.g. on
 * one stream).  Nothing of an earlier call's scratch contents is used.
 *
 * No counterpart in the reference.  For callers that cannot allocate: code
 * under HIP graph capture, tight serving loops, harnesses that forbid it.
 */
#define ATTGNN_NOALLOC_SCRATCH_ELEMS (16 * 1024 * 1024)
int attgnn_spmm_block_noalloc(const int32_t *rp, const int32_t *col, const void *val, int tile_type,
                              const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                              const void *X, int x_type, int64_t n_x, void *Y, int y_type, int64_t m_ori,
                              int64_t dim_n, int window, int wide, int mode, const float *row_scale,
                              const float *col_scale, int flags, attgnn_stream_t stream);

/*
 * attgnn_spmm_block over compact tiles (attgnn_blocks_compact_export: row masks,
 * block offsets, packed values) of a structure built with wide = 32; same
 * arguments otherwise, same results bit for bit (the multiply-adds are the
 * same, absent entries contribute exact zeros in both).
 *
 * Replaces SpMM_sr_bcrs.forward_*_gnn (gcn/SpMM/src/benchmark.cpp:35-130) on
 * a smaller tile format; no reference counterpart for the format itself.
 */
int attgnn_spmm_compact(const int32_t *rp, const int32_t *col, const uint16_t *masks,
                        const int64_t *block_base, const void *values, int tile_type,
                        const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                        const void *X, int x_type, int64_t n_x, void *Y, int y_type, int64_t m_ori,
                        int64_t dim_n, int window, int mode, const float *row_scale,
                        const float *col_scale, float *workspace, int64_t workspace_elems, int flags,
                        attgnn_stream_t stream);

/*
 * Striped node order and the owner-split SpMM.  No reference counterpart: the
 * reference gathers X rows in the graph's own order
 * (gcn/SpMM/src/spmmKernel.cu:158-265); these entry points compute the same
 * product as attgnn_spmm_compact on a relabelled graph.
 *
 * With owners = 2^L (2 to 32) and m nodes (m a multiple of 64 * owners,
 * q = m / owners), node j sits at
 *     stripe_pos(j) = ((j >> 6) mod owners) * q + (j >> (6 + L)) * 64 + (j mod 64),
 * a bijection on [0, m): 64-node stripes dealt round-robin to the owners, each
 * owner's stripes one contiguous range of q positions.  A striped structure
 * is built (attgnn_blocks_build / _device, compact tiles) from the CSR whose
 * rows stay in the graph's order and whose column ids are replaced by their
 * stripe positions (re-sorted within each row); its parts' columns are then
 * ascending, so each part's columns fall into consecutive owner segments.
 *
 * attgnn_stripe_segments_device: owner_seg[(owners + 1) p + o], o = 0..owners,
 * the first position in part p's real columns [rp[2p], rp[2p+1]) whose id is
 * >= o * q (owner_seg[.. + owners] = rp[2p+1]); parts with window < 0 get
 * rp[2p] throughout.  Device int32, (owners + 1) * parts entries.
 *
 * attgnn_spmm_compact_striped: Y' = A X' over a striped structure.  X' [n_x >= m]
 * [dim_n] holds node j's features at row stripe_pos(j); Y' [m][dim_n] (y_type)
 * receives window row j at stripe_pos(j).  With F16 or BF16 tiles the parts
 * are split by owner: with S = ceil(dim_n / (16 T)) feature slices (T =
 * features per lane, 8 for 16-bit data at dim_n >= 128) and G owner groups
 * (a power of two up to min(owners, 8 / S): by default the fewest that leave
 * each XCD about 16 MB of X), each part is taken by G * S tasks, at most one per
 * XCD; the task on XCD slot (gi, s) takes slice s of the segments of owners
 * gi, gi + G, ..., so each XCD's L2 caches its owners' rows of one slice.
 * Every task stores its fp32 tile in the workspace and a second kernel sums
 * each window's tiles in (part, group) order: deterministic, and independent of
 * the structure's partSize when the parts are a canonical task table
 * (attgnn_blocks_tasks_device).  Otherwise (or with
 * ATTGNN_SPMM_NO_XCD_SLICES) the tasks map as in attgnn_spmm_compact and
 * only the row positions change.  The split also needs dim_n to be a multiple
 * of 16 T, X to span fewer than 2^31 bytes and rows of fewer than 2^24 bytes
 * (32-bit buffer offsets); otherwise the usual mapping runs.  mode NATIVE
 * (F16 / BF16), BF16X3 or F32; no FP8.  workspace:
 * attgnn_spmm_striped_workspace_elems floats.  flags as for
 * attgnn_spmm_compact (the k-loop schedule bits apply to the usual mapping
 * only; the split has one schedule) plus the owner-group bits 12-14.
 * ATTGNN_E_ARG for owners not a power of two in [2, 32], m not a
 * multiple of 64 * owners, n_x < m, or a missing owner_seg where the split
 * runs.
 */
int attgnn_stripe_segments_device(const int32_t *rp, const int32_t *window_of_part, const int32_t *col,
                                  int64_t parts, int owners, int64_t m, int32_t *owner_seg,
                                  attgnn_stream_t stream);
int64_t attgnn_spmm_striped_workspace_elems(int64_t parts, int window, int64_t dim_n, int owners);
int attgnn_spmm_compact_striped(const int32_t *rp, const int32_t *col, const uint16_t *masks,
                                const int64_t *block_base, const void *values, int tile_type,
                                const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                                const int32_t *owner_seg, int owners, const void *X, int x_type, int64_t n_x,
                                void *Y, int y_type, int64_t m, int64_t dim_n, int window, int mode,
                                float *workspace, int64_t workspace_elems, int flags, attgnn_stream_t stream);

/*
 * Edge-map SpMM: Y_h = A_h . X_h with A_h[i, j] = edge_val[h*nnz + map[slot(i,j)]],
 * tiles given by the edge-id map of attgnn_blocks_build(ATTGNN_SRC_MAP) (or its
 * transposed map, which aggregates with A^T on a symmetric pattern).
 *
 * Replaces ATT_SpMM_v{1,3}.spmm_{fp16,tf32}_h_m  (ATT_SpMM_v3.spmm_fp16_h_m, caller
 *            end2end_head/gat_final/mgat_csr/mgat_conv.py:99-106; known answer
 *            benchmark-v3/spmm-benchmark/test/spmm_fp16/test_spmm_h_m.py:45-66)
 *          ATT_SpMM_v{1,3}.spmm_{fp16,tf32}_m_h  (mgat_conv.py:79-86;
 *            test_spmm_m_h.py:49-84)
 *
 * edge_val[heads][nnz] (F16/BF16/F32, matching x_type), X in x_layout
 * (HMD [heads][n_x][dim_n] or MHD [n_x][heads][dim_n]), Y of y_type (F32 /
 * F16 / BF16, FP32 accumulation rounded once) in out_layout (HMD:
 * [heads][m_ori][dim_n], MHD: [m_ori][heads][dim_n]).  flags: 0 or
 * ATTGNN_SPMM_ROWS_COVERED (see attgnn_spmm_block).
 * mode: NATIVE for F16/BF16; BF16X3 or F32 for F32.  workspace: at least
 * heads * attgnn_spmm_workspace_elems(parts, window, dim_n) floats.
 */
int attgnn_spmm_edge_map(const int32_t *rp, const int32_t *col, const int32_t *map,
                         const void *edge_val, int val_type, int64_t nnz,
                         const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                         const void *X, int x_type, int x_layout, int64_t heads, int64_t n_x,
                         void *Y, int y_type, int64_t m_ori, int64_t dim_n, int out_layout, int window,
                         int wide, int mode, float *workspace, int64_t workspace_elems, int flags,
                         attgnn_stream_t stream);

/*
 * attgnn_spmm_edge_map over the compact form of the edge-id map (wide = 32,
 * window <= 16): masks, block_base and ids as written by
 * attgnn_blocks_compact_export / attgnn_blocks_compact_fill_device for the
 * map (tile type I32).  A block reads its row masks and its edges' ids -- a
 * few hundred bytes on sparse windows -- instead of the window x 32 int32 map
 * tile.  Same arguments otherwise; bitwise equal to attgnn_spmm_edge_map.
 * Replaces the same reference entry points as attgnn_spmm_edge_map (the
 * compact form itself has no counterpart in the reference).
 */
int attgnn_spmm_edge_map_compact(const int32_t *rp, const int32_t *col, const uint16_t *masks,
                                 const int64_t *block_base, const int32_t *ids, const void *edge_val,
                                 int val_type, int64_t nnz, const int32_t *window_of_part, const int32_t *atomic,
                                 int64_t parts, const void *X, int x_type, int x_layout, int64_t heads,
                                 int64_t n_x, void *Y, int y_type, int64_t m_ori, int64_t dim_n, int out_layout,
                                 int window, int mode, float *workspace, int64_t workspace_elems, int flags,
                                 attgnn_stream_t stream);

/*
 * attgnn_spmm_edge_map_compact fused with the SDDMM over the same structure:
 * besides Y, every edge e in the structure at window row r and column c gets
 *   dot_out[h*nnz + e] = sum_f dot_x[h, r, f] * X[h, c, f]   (fp32),
 * the products taken from the X rows the aggregation gathers anyway (dot_x
 * has X's type and layout, dot_rows rows, 16-byte aligned).  Run over the
 * structure of A^T with A's edge ids (GAT backward: X = dL/dh, dot_x = the
 * projected features) it gives the aggregation gradient and the
 * edge-probability gradient <dh[i], Xt[j]> in A's edge order in one pass over
 * the gathered rows; bitwise equal to attgnn_spmm_edge_map_compact plus
 * attgnn_sddmm_edges over A.  F16 / BF16 only (native mode), dim_n 32, 64 or
 * 128; ATTGNN_E_UNSUPPORTED otherwise.  The k-loop schedule flags are ignored.
 *
 * Replaces the pair ATT_SpMM_v{1,3}.spmm_*_h_m (transposed map) +
 * ATT_SDDMM_v{1,3}.sddmm_*_h_nnz of the attention backward
 *   (end2end_head/gat_final/mgat_csr/mgat_conv.py:91-151).
 */
int attgnn_spmm_edge_map_compact_dot(const int32_t *rp, const int32_t *col, const uint16_t *masks,
                                     const int64_t *block_base, const int32_t *ids, const void *edge_val,
                                     int val_type, int64_t nnz, const int32_t *window_of_part,
                                     const int32_t *atomic, int64_t parts, const void *X, int x_type,
                                     int x_layout, int64_t heads, int64_t n_x, void *Y, int y_type,
                                     int64_t m_ori, int64_t dim_n, int out_layout, int window,
                                     const void *dot_x, int64_t dot_rows, float *dot_out,
                                     float *workspace, int64_t workspace_elems, int flags,
                                     attgnn_stream_t stream);

/*
 * Attention aggregation with the probabilities rebuilt in the kernel:
 * Y_h = P_h . X_h over the edge-id map's tiles, where the entry of edge e =
 * (r, c) (e = map[slot], A's CSR position) is
 *   P_h[r, c] = exp(LeakyReLU_slope(row_logit(h, r) + col_logit(h, c))
 *                   - row_offset(h, r) - col_offset(h, c)),
 * then dropout with drop_p from the same (seed, h, e) hash as
 * attgnn_edge_softmax_fwd, kept values scaled by 1/(1-drop_p); NULL offsets
 * count as 0.  Element (h, v) of the logit terms is at
 * [h*logit_head_stride + v*logit_node_stride], of the offsets at
 * [h*offset_head_stride + v*offset_node_stride].  With the map of A,
 * row_logit = s0, col_logit = s1, row_offset = the softmax's lse, this equals
 * attgnn_spmm_edge_map over the p written by attgnn_edge_softmax_fwd bit for
 * bit (p in X's type); with the transposed map, row_logit = s1, col_logit =
 * s0, col_offset = lse it is the A^T aggregation of the backward -- and p is
 * never stored.  Other arguments as attgnn_spmm_edge_map (X, Y, layouts,
 * mode, workspace, flags).
 *
 * Replaces the aggregation calls ATT_SpMM_v{1,3}.spmm_*_{h_m,m_h} of
 *   end2end_head/gat_final/mgat_csr/mgat_conv.py:79-86, 99-106 together with
 *   the probability tensor they read (:180-197).
 */
typedef struct attgnn_attn_terms {
    const float *row_logit, *col_logit;
    const float *row_offset, *col_offset;  /* may be NULL */
    int64_t logit_head_stride, logit_node_stride;
    int64_t offset_head_stride, offset_node_stride;
    int64_t n_nodes;                       /* terms cover nodes [0, n_nodes) */
    float slope, drop_p;
    uint64_t seed;
} attgnn_attn_terms;
int attgnn_spmm_attention(const int32_t *rp, const int32_t *col, const int32_t *map,
                          const int32_t *window_of_part, const int32_t *atomic, int64_t parts,
                          const attgnn_attn_terms *terms, const void *X, int x_type, int x_layout,
                          int64_t heads, int64_t n_x, void *Y, int y_type, int64_t m_ori, int64_t dim_n,
                          int out_layout, int window, int wide, int mode, float *workspace,
                          int64_t workspace_elems, int flags, attgnn_stream_t stream);

/*
 * Row sums of per-edge values through the edge-id map:
 * out(h, i) = sum_{slots of row i} edge_val[h*nnz + map[slot]] (fp32).
 *
 * Replaces ATT_SpMM_v{1,3}.spmm_{fp16,tf32}_{h_m,m_h}_rowsum
 *            (callers end2end_head/gat_final/mgat_csr/mgat_conv.py:42-59,126-132;
 *             known answer test_spmm_h_m_ones.py:45-70)
 * out is [heads][m_rows] (HMD) or [m_rows][heads] (MHD); rows no window
 * covers are zero.  Deterministic (fixed-order reduction per window).
 */
int attgnn_rowsum_edge_map(const int32_t *rp, const int32_t *map, const int32_t *window_of_part,
                           int64_t parts, const void *edge_val, int val_type, int64_t nnz,
                           int64_t heads, float *out, int out_layout, int64_t m_rows, int window,
                           int wide, attgnn_stream_t stream);

/* ====================================================================== */
/* Window-block SDDMM and per-edge attention terms (device)                */
/* ====================================================================== */

/*
 * out[h*nnz + e] = < A[h, i, :], B[h, j, :] > for every edge e = (i, j) in the
 * edge-id map; A element (h,i,d) at A[h*a_head_stride + i*a_row_stride + d].
 *
 * Replaces ATT_SDDMM_v{1,3}.sddmm_{fp16,tf32}_h_nnz
 *            (benchmark-v3/sddmm-benchmark/include/compute_utils.h:142-280,427-530,
 *             output_tile.h:5-84; caller end2end_head/gat_final/mgat_csr/mgat_conv.py:109-117;
 *             known answer test/sddmm_fp16/test_sddmm_fp16.py:50-69)
 * in_type F16/BF16 (MFMA in that type) or F32 (mode BF16X3 or F32);
 * out_type F16/BF16/F32.  Edges of the map never visited are left untouched.
 */
int attgnn_sddmm_block(const int32_t *rp, const int32_t *col, const int32_t *map,
                       const int32_t *window_of_part, int64_t parts, const void *A,
                       int64_t a_head_stride, int64_t a_row_stride, const void *B,
                       int64_t b_head_stride, int64_t b_row_stride, int in_type,
                       int64_t heads, int64_t n_rows, int64_t dim, void *out, int out_type,
                       int64_t nnz, int window, int wide, int mode, attgnn_stream_t stream);

/*
 * attgnn_sddmm_block over a compact edge list (attgnn_blocks_compact_export of
 * an edge-id map built with wide = 32: block_base, edge ids, positions): all
 * heads in one pass, the dense map is not read.  Same results.
 *
 * Replaces ATT_SDDMM_v{1,3}.sddmm_{fp16,tf32}_h_nnz (see attgnn_sddmm_block).
 */
int attgnn_sddmm_edges(const int32_t *rp, const int32_t *col, const int64_t *block_base,
                       const int32_t *edge_ids, const uint16_t *positions,
                       const int32_t *window_of_part, int64_t parts, const void *A,
                       int64_t a_head_stride, int64_t a_row_stride, const void *B,
                       int64_t b_head_stride, int64_t b_row_stride, int in_type, int64_t heads,
                       int64_t n_rows, int64_t dim, void *out, int out_type, int64_t nnz, int window,
                       int mode, attgnn_stream_t stream);

/*
 * scores[h*s_head_stride + i*s_row_stride] =
 *     sum_d X[h*x_head_stride + i*x_row_stride + d] * vec[h*dim + d]
 * (fp32 accumulate) -- the per-node half of the attention projections.
 * (s_head_stride, s_row_stride) = (n_rows, 1) gives [H, N] scores, (1, heads)
 * node-major [N, H] scores, the layout the edge softmax reads fastest.
 *
 * Part of ATT_SDDMM_v{1,3}.att_{fp16,tf32}_h_nnz / attv2_{fp16,tf32}_h_nnz
 *            (compute_utils.h:1144-1218; callers end2end_head/gat_final/mgat_csr/mgat_conv.py:21-30,
 *             end2end_head/gatv2_final/mgat_csr/mgat_conv.py:22-31)
 */
int attgnn_node_scores(const void *X, int x_type, int64_t x_head_stride, int64_t x_row_stride,
                       const void *vec, int vec_type, int64_t heads, int64_t n_rows, int64_t dim,
                       float *scores, int64_t s_head_stride, int64_t s_row_stride, attgnn_stream_t stream);

/*
 * attgnn_node_scores for two vectors over the same X in one pass:
 * scores0[h, i] = X_h[i] . vec0_h and scores1[h, i] = X_h[i] . vec1_h, each
 * bitwise equal to its attgnn_node_scores call (GAT, where both attention
 * terms read the same projected features).
 *
 * Replaces the pair of ATT_SDDMM_v{1,3}.att_*_h_nnz node terms
 *   (end2end_head/gat_final/mgat_csr/mgat_conv.py:20-31).
 */
int attgnn_node_scores_pair(const void *X, int x_type, int64_t x_head_stride, int64_t x_row_stride,
                            const void *vec0, const void *vec1, int vec_type, int64_t heads, int64_t n_rows,
                            int64_t dim, float *scores0, float *scores1, int64_t s_head_stride,
                            int64_t s_row_stride, attgnn_stream_t stream);

/*
 * out[h*nnz + e] = S0(h, i) + S1(h, j) for every edge e = (i, j) of the map, with
 * S0(h,i) = s0[h*s0_head_stride + i*s0_row_stride] (or the typed g0 at the same
 * strides when g0 != NULL) and S1 likewise from s1 (NULL = 0).
 *
 * Replaces the per-edge half of ATT_SDDMM_v{1,3}.att_* / attv2_* (output_tile.h:89-202)
 *          ATT_SDDMM_v{1,3}.sddmm_{fp16,tf32}_ones_h_nnz (g0 = g, s1 = NULL;
 *            caller end2end_head/gat_final/mgat_csr/mgat_conv.py:142-149;
 *            known answer test/sddmm_fp16/test_sddmm_fp16_ones.py:51-68)
 */
int attgnn_edge_from_nodes(const int32_t *rp, const int32_t *col, const int32_t *map,
                           const int32_t *window_of_part, int64_t parts, const float *s0,
                           int64_t s0_head_stride, int64_t s0_row_stride, const float *s1,
                           int64_t s1_head_stride, int64_t s1_row_stride, const void *g0,
                           int g0_type, int64_t heads, int64_t n_rows, void *out, int out_type,
                           int64_t nnz, int window, int wide, attgnn_stream_t stream);

/*
 * Backward of the per-node attention terms s0 = Xs.a0 and s1 = Xt.a1, fused
 * with the aggregation gradient, over [n_rows][heads][dim] features (x_type
 * F32 / F16 / BF16, contiguous) with a0, a1 [heads][dim] fp32 and
 * ds0, ds1 [heads][n_rows] fp32 (dL/ds0, dL/ds1):
 *   dxs[i,h,:] = ds0[h,i] a0[h,:]
 *   dxt[i,h,:] = base[i,h,:] + ds1[h,i] a1[h,:]       (base fp32, NULL = 0)
 *   da0[h,:]   = sum_i ds0[h,i] xs[i,h,:],  da1[h,:] = sum_i ds1[h,i] xt[i,h,:]
 * dxs == NULL (only when xs == xt, as in GAT where both are X W): the two
 * gradients are summed into dxt.  dxs / dxt in x_type, rounded once;
 * da0 / da1 fp32, reduced in a fixed order (deterministic).
 * workspace: attgnn_attention_node_bwd_workspace_elems() floats.
 *
 * Replaces the attention-vector and feature gradients of the reference's
 * backward (end2end_head/gat_final/mgat_csr/mgat_conv.py:34-69, 91-151,
 * where a0 is paired with column sums and cross terms are added), with the
 * exact gradient.
 */
int64_t attgnn_attention_node_bwd_workspace_elems(int64_t n_rows, int64_t heads, int64_t dim);
int attgnn_attention_node_bwd(const void *xs, const void *xt, int x_type, int64_t n_rows, int64_t heads,
                              int64_t dim, const float *a0, const float *a1, const float *ds0, const float *ds1,
                              const float *base, void *dxs, void *dxt, float *da0, float *da1, float *workspace,
                              int64_t workspace_elems, attgnn_stream_t stream);

/*
 * Segments of a CSR's rows for the per-edge attention kernels: every row is
 * cut into pieces of at most 32 consecutive entries, so a long row becomes
 * many short tasks.  Segment s covers entries [seg_ptr[s], seg_ptr[s+1]) of
 * row seg_row[s]; the segments of row i are [row_seg[i], row_seg[i+1]) (none
 * for an empty row); seg_slot[s] = -1 when s is its row's only segment, else
 * its index 0..n_slots-1 among the segments of rows that have several.  All
 * arrays int32 in device memory; att_amd.graphs.csr_segments builds them.
 */
typedef struct attgnn_segments {
    const int32_t *seg_ptr;   /* [n_seg + 1] */
    const int32_t *seg_row;   /* [n_seg]     */
    const int32_t *seg_slot;  /* [n_seg]     */
    const int32_t *row_seg;   /* [rows + 1]  */
    int64_t n_seg, n_slots;
} attgnn_segments;

/*
 * GATv2 attention as the GATv2 paper defines it, over CSR rows of A:
 *   z[h*nnz + e] = sum_d att[h,d] LeakyReLU_slope(x0[i,h,d] + x1[j,h,d]), e = (i, j)
 *   lse[h*n + i] = log sum_{e in row i} exp(z),  p[h*nnz + e] = exp(z - lse)
 * with dropout on p as attgnn_edge_softmax_fwd (drop_p, seed).  x0, x1
 * [n][heads][dim] of x_type (F32 / F16 / BF16, contiguous), att fp32
 * [heads][dim]; dim in {16, 32, 64, 128} or any dim <= 128.  row_segments:
 * the segments of A's rows.
 *
 * No counterpart in the reference: its GATv2 (end2end_ori/gatv2_final/mgat_csr/mgat_conv.py:174-196,
 * ATT_SDDMM_v*.attv2_*) applies the LeakyReLU before a separable dot product.
 */
int attgnn_gatv2_fwd(const int32_t *rowptr, const int32_t *col, const attgnn_segments *row_segments, int64_t n,
                     int64_t nnz, int64_t heads, int64_t dim, const void *x0, const void *x1, int x_type,
                     const float *att, float slope, float drop_p, uint64_t seed, float *z, float *lse, void *p,
                     int p_type, attgnn_stream_t stream);

/*
 * Backward of attgnn_gatv2_fwd given dp = dL/dp (after dropout, dp_type):
 *   dz[h*nnz + e]  (softmax backward, fp32)
 *   dx0[i,h,:] = sum_{e in row i} dz_e att[h,:] * LeakyReLU'(x0[i] + x1[j])
 *   dx1[j,h,:] = base[j,h,:] + sum_{e = (i, j)} dz_e att[h,:] * LeakyReLU'(x0[i] + x1[j])
 *   datt[h,:]  = sum_e dz_e LeakyReLU(x0[i] + x1[j])   (fixed-order reduction)
 * The column sums walk the CSR of A^T given by col_segments (segments of A^T's
 * rows), t_col (row of A of each A^T entry) and t_perm (A edge id of each A^T
 * entry, int64).  base fp32 [n][heads][dim] or NULL; dx0 / dx1 of x_type.
 * workspace: attgnn_gatv2_bwd_workspace_elems() floats.  Deterministic.
 *
 * No counterpart in the reference (see attgnn_gatv2_fwd).
 */
int64_t attgnn_gatv2_bwd_workspace_elems(int64_t heads, int64_t dim, const attgnn_segments *row_segments,
                                         const attgnn_segments *col_segments);
int attgnn_gatv2_bwd(const int32_t *rowptr, const int32_t *col, const attgnn_segments *row_segments,
                     const int32_t *t_col, const int64_t *t_perm, const attgnn_segments *col_segments, int64_t n,
                     int64_t nnz, int64_t heads, int64_t dim, const void *x0, const void *x1, int x_type,
                     const float *att, float slope, float drop_p, uint64_t seed, const float *z, const float *lse,
                     const void *dp, int dp_type, const float *base, float *dz, void *dx0, void *dx1, float *datt,
                     float *workspace, int64_t workspace_elems, attgnn_stream_t stream);

/* ====================================================================== */
/* Fused edge softmax over CSR rows (device)                               */
/* ====================================================================== */

/*
 * Long rows.  The row kernels below give each row a 16-lane group, so on a
 * power-law graph a hub row of a few thousand entries outlasts all the others.
 * Every row entry point takes an optional list of the rows with more than
 * ATTGNN_LONG_ROW entries (long_rows, n_long; NULL / 0 = none): the group
 * kernel then skips those rows and a 256-thread workgroup per (listed row,
 * head) takes them.  The list must hold every row above the threshold (any
 * order; listed rows at or below it are ignored).  attgnn_long_rows fills it
 * from a host copy of the row pointers and returns its length (rows = NULL:
 * count only; -1 on bad arguments).  Results depend on the list only through
 * the summation order of the long rows, and are deterministic either way.
 */
#define ATTGNN_LONG_ROW 256
int64_t attgnn_long_rows(const int32_t *rowptr_host, int64_t n, int32_t *rows);

/*
 * Attention probabilities of one layer, fused:
 *   z = s0[h,i] + s1[h,j];  v = LeakyReLU_slope(z);  p = exp(v - lse_i) with
 *   lse_i = log sum_j exp(v) (max subtracted, FP32);  dropout with probability
 *   drop_p from a counter-based hash of (seed, h, e), kept values scaled by
 *   1/(1-drop_p).  Writes p[h*nnz + e] (p_type) and lse[h*n + i] (fp32);
 *   p = NULL writes the log-sum-exp only (for attgnn_spmm_attention, which
 *   rebuilds p where it is used).
 * s0/s1 element (h, i) at s[h*s_head_stride + i*s_row_stride].
 *
 * Replaces the LeakyReLU / exp / rowsum / dropout / divide chain of
 *   end2end_head/gat_final/mgat_csr/mgat_conv.py:180-197 (exp without max
 *   subtraction there; the normalised result is the same) -- slope 0.2 for GAT,
 *   slope 1 for the reference's GATv2 logits
 *   (end2end_ori/gatv2_final/mgat_csr/mgat_conv.py:179-196).
 */
int attgnn_edge_softmax_fwd(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz,
                            int64_t heads, const float *s0, const float *s1, int64_t s_head_stride,
                            int64_t s_row_stride, float slope, float drop_p, uint64_t seed, void *p,
                            int p_type, float *lse, const int32_t *long_rows, int64_t n_long,
                            attgnn_stream_t stream);

/*
 * Backward of attgnn_edge_softmax_fwd: p is recomputed from (s0, s1, lse) and
 * the dropout mask regenerated from (seed, h, e); dp = dL/d(dropped p).
 * Writes dz[h*nnz + e] = dL/dz and ds0[h*n + i] = sum_j dz[h, (i, j)].
 *
 * Replaces the backward of the same chain (mgat_conv.py:34-69, 91-151), with
 * the exact gradient.
 */
int attgnn_edge_softmax_bwd(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz,
                            int64_t heads, const float *s0, const float *s1, int64_t s_head_stride,
                            int64_t s_row_stride, const float *lse, const void *dp, int dp_type,
                            float slope, float drop_p, uint64_t seed, float *dz, float *ds0,
                            const int32_t *long_rows, int64_t n_long, attgnn_stream_t stream);

/*
 * Backward of attgnn_edge_softmax_fwd reduced to the two node gradients the
 * attention layers need: ds0[h*n + i] = sum_j dz[h, (i, j)] (row sums) and
 * ds1[h*n + j] = sum_i dz[h, (i, j)] (column sums, through the transpose:
 * t_rowptr [n+1] holds the rows of A^T and t_perm[q] the CSR position of A^T
 * entry q, as attgnn_csr_transpose_device writes them; long_rows / t_long_rows
 * list the long rows of A and of A^T).  dz is kept in the
 * workspace only (fp32, attgnn_edge_softmax_bwd_sums_workspace_elems
 * elements), edge-major for 1, 2, 4 or 8 heads.  Results are bitwise equal to
 * attgnn_edge_softmax_bwd followed by attgnn_gather_rowsum.
 *
 * Replaces the backward of the softmax chain together with the column-sum
 * call ATT_SpMM_v{1,3}.spmm_*_h_m_rowsum over the transposed map
 *   (end2end_head/gat_final/mgat_csr/mgat_conv.py:42-48, 91-151).
 */
int64_t attgnn_edge_softmax_bwd_sums_workspace_elems(int64_t nnz, int64_t heads);
int attgnn_edge_softmax_bwd_sums(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz,
                                 int64_t heads, const float *s0, const float *s1, int64_t s_head_stride,
                                 int64_t s_row_stride, const float *lse, const void *dp, int dp_type,
                                 float slope, float drop_p, uint64_t seed, const int32_t *t_rowptr,
                                 const int64_t *t_perm, const int32_t *long_rows, int64_t n_long,
                                 const int32_t *t_long_rows, int64_t n_t_long, float *ds0, float *ds1,
                                 float *workspace, int64_t workspace_elems, attgnn_stream_t stream);

/*
 * out[h*n + j] = sum over CSR row j of vals[h*nnz + perm[e]] (perm[e] = -1 skips):
 * with perm the CSR position of each entry's transpose this is a column sum.
 *
 * Replaces ATT_SpMM_v{1,3}.spmm_*_h_m_rowsum called with the transposed map
 *   (end2end_head/gat_final/mgat_csr/mgat_conv.py:42-48).
 */
int attgnn_gather_rowsum(const int32_t *rowptr, const int64_t *perm, int64_t n, int64_t nnz,
                         int64_t heads, const float *vals, float *out, const int32_t *long_rows,
                         int64_t n_long, attgnn_stream_t stream);

/*
 * Per-head range of the logits z[h, e=(i,j)] = s0[h,i] + s1[h,j] over the
 * CSR's nnz edges (s0 / s1 strided as above): range[3h..3h+2] = (min, max,
 * 1/(max - min)), the last 0 when max == min or there are no edges; where[4h..
 * 4h+3] = (row, column) of the first edge in CSR order attaining the minimum,
 * then of the one attaining the maximum (-1 without edges).  workspace: a
 * device buffer of attgnn_edge_score_range_workspace_elems(heads) 64-bit words
 * (per-workgroup partial minima and maxima, overwritten).  long_rows /
 * n_long: the list of attgnn_long_rows (NULL / 0 = none), whose rows then get
 * a workgroup each, as in the softmax.  Deterministic (integer min / max
 * reductions, no atomics).  heads <= 256, nnz < 2^31.
 *
 * Replaces the per-head torch.max / torch.min over the logits and the
 * division by their difference of the multi-head GATv2 layer
 *   (end2end_head/gatv2_final/mgat_csr/mgat_conv.py:184-186); the layer scales
 *   s0 and s1 by range[3h+2] before attgnn_edge_softmax_fwd (the softmax is
 *   shift-invariant, so subtracting the minimum changes nothing).
 */
#define ATTGNN_RANGE_BLOCKS 8192
int64_t attgnn_edge_score_range_workspace_elems(int64_t heads);
int attgnn_edge_score_range(const int32_t *rowptr, const int32_t *col, int64_t n, int64_t nnz, int64_t heads,
                            const float *s0, const float *s1, int64_t s_head_stride, int64_t s_row_stride,
                            unsigned long long *workspace, int64_t workspace_elems, float *range, int32_t *where,
                            const int32_t *long_rows, int64_t n_long, attgnn_stream_t stream);

/*
 * Gradients of the raw node terms of the min-max logit map from those of the
 * scaled terms.  With c_h = range[3h+2] (attgnn_edge_score_range) the layer
 * feeds s0 c and s1 c to the softmax; given ds0s / ds1s = dL/d(s0 c), dL/d(s1 c)
 * (fp32, [heads][n] contiguous) this writes ds0 = c ds0s, ds1 = c ds1s, then,
 * with U_h = sum_i s0[h,i] ds0s[h,i] + s1[h,i] ds1s[h,i] and k_h = U_h c_h^2,
 * subtracts k_h at the row (ds0) and column (ds1) of the maximum's edge and
 * adds k_h at those of the minimum's (where[4h..4h+3]; in that order, so an
 * element hit twice gets (v - k) + k).  s0 / s1 strided like the scores of
 * attgnn_edge_score_range.  workspace: attgnn_minmax_grads_workspace_elems(
 * heads) floats.  Deterministic (fixed-order sums, no atomics).
 *
 * Replaces the autograd of (z - min) / (max - min) through torch.min / max of
 * the multi-head GATv2 layer (end2end_head/gatv2_final/mgat_csr/mgat_conv.py:184-186).
 */
int64_t attgnn_minmax_grads_workspace_elems(int64_t heads);
int attgnn_minmax_score_grads(const float *s0, const float *s1, int64_t n, int64_t heads, int64_t s_head_stride,
                              int64_t s_row_stride, const float *range, const int32_t *where, const float *ds0s,
                              const float *ds1s, float *ds0, float *ds1, float *workspace,
                              int64_t workspace_elems, attgnn_stream_t stream);

/* ====================================================================== */
/* Block-diagonal packing for batched inference (device)                   */
/* ====================================================================== */

/*
 * Copy segments described by a device-resident table of nsegs entries
 * (attgnn_pack_segment_bytes() bytes each: const void *src; void *dst;
 * int64_t count; int64_t first; int32_t offset; int32_t kind), one workgroup
 * per chunk of chunk_bytes (a multiple of 16): chunk c copies the part of
 * segment chunk_seg[c] from item (c - first) * chunk_bytes / item size;
 * chunk_seg[c] = -1 copies nothing.  kind 0: int32 dst[i] = src[i] < 0 ?
 * src[i] : src[i] + offset; kind 1: int32 dst[i] = (src ? src[i] : 0) + offset
 * (src NULL = fill); kind 2: copy count bytes; kind 3: int64 dst[i] =
 * (src ? src[i] : 0) + offset (compact block offsets).
 * One launch packs a whole batch of cached graphs into one block-diagonal
 * window-block structure; the tables live in device memory so the launch can
 * be captured in a HIP graph (fixed n_chunks, unused chunks -1) and replayed
 * with new tables.
 *
 * No counterpart in the reference (it runs one graph per launch).
 */
int attgnn_pack_segments(const void *segs, int64_t nsegs, const int32_t *chunk_seg, int64_t n_chunks,
                         int64_t chunk_bytes, attgnn_stream_t stream);
int64_t attgnn_pack_segment_bytes(void);

/* ====================================================================== */
/* Training loss                                                           */
/* ====================================================================== */

/*
 * Mean negative log-likelihood of a row-wise log-softmax and its gradient in
 * one pass (replaces F.log_softmax + F.nll_loss of the reference's GCN,
 * end2end_ori/gcn_final/mgat_csr/gat_mgnn.py:39,66).  logits: [n_rows][ld]
 * of logits_type (F32 / F16 / BF16), the first `classes` columns of each row
 * being the classes; labels: int64 [n_rows]; rows: the n_sel row indices the
 * mean runs over (NULL: rows 0 .. n_sel-1).  loss (may be NULL): loss[0] =
 * (1/n_sel) sum over the selected rows r of (logsumexp(logits[r, :classes])
 * - logits[r, labels[r]]), summed in a fixed order (the same bits run to
 * run), with workspace of attgnn_nll_log_softmax_workspace_elems(n_sel)
 * floats.  grad (may be NULL, not both): for every selected row and class c,
 * grad[r][c] = g (softmax(logits[r, :classes])[c] - [c == labels[r]]) / n_sel
 * in logits_type, g = grad_scale[0] (a device float; NULL: 1) -- the
 * backward of the mean scaled by the upstream gradient; the other elements
 * of grad are not written.  FP32 arithmetic.  A label outside [0, classes)
 * gives NaN for its row (loss and gradient), a row index outside [0, n_rows)
 * NaN in the loss (nothing stored).
 */
int64_t attgnn_nll_log_softmax_workspace_elems(int64_t n_sel);
int attgnn_nll_log_softmax(const void *logits, int logits_type, int64_t n_rows, int64_t ld, int classes,
                           const int64_t *labels, const int64_t *rows, int64_t n_sel, float *loss, void *grad,
                           const float *grad_scale, float *workspace, int64_t workspace_elems,
                           attgnn_stream_t stream);

/*
 * ReLU + dropout in one pass (replaces F.relu and F.dropout between the
 * reference GCN's layers, end2end_ori/gcn_final/mgat_csr/gat_mgnn.py:17-40).
 * Forward: y[i] = x[i] / (1 - p) if x[i] > 0 (or NaN) and element i is kept,
 * else 0;
 * element i is kept when the top 24 bits of lowbias32(lowbias32(seed_lo ^
 * seed_hi) + i) are at least p * 2^24 (lowbias32: x ^= x >> 16;
 * x *= 0x7feb352d; x ^= x >> 15; x *= 0x846ca68b; x ^= x >> 16; i taken
 * mod 2^32), so the mask depends on (seed, i) only; bit b of mask[j] records
 * whether element 8j + b passed (kept and positive).  Backward: dx[i] = dy[i] / (1 - p) where the
 * bit is set, else 0.  n (elements) a multiple of 8, x / y / dy / dx
 * 16-byte aligned, type F32 / F16 / BF16, p in [0, 1]; mask holds n / 8
 * bytes.  No counterpart kernel in the reference (torch ops there).
 */
int attgnn_relu_dropout_fwd(const void *x, int type, int64_t n, float p, uint64_t seed, void *y, uint8_t *mask,
                            attgnn_stream_t stream);
int attgnn_relu_dropout_bwd(const void *dy, int type, int64_t n, float p, const uint8_t *mask, void *dx,
                            attgnn_stream_t stream);

/*
 * Host staging of one serving graph (no counterpart in the reference, which
 * serves nothing).  attgnn_serve_stage_layout: the byte offsets, in one
 * buffer, of the CSR row pointers padded to whole windows (int32
 * [n_pad + 1], padding rows empty; offsets[0]), the column ids (int32 [nnz];
 * offsets[1]), the degrees d_i = 1 / sqrt(row count) (float [n_pad], 0 for an
 * empty row; offsets[2]) and the features (bf16 [n_pad][feat], padding rows
 * zero; offsets[3]), each 256-byte aligned, with n_pad = n rounded up to a
 * multiple of window (window for n = 0) in offsets[4]; returns the total
 * bytes, or ATTGNN_E_ARG.
 * attgnn_serve_stage_host: writes that buffer into the host memory `out` (at
 * least the total bytes) from a CSR (indptr int64 [n + 1], indptr[0] = 0,
 * indptr[n] = nnz; indices int32 [nnz]) and float features [n][feat]
 * (rounded to nearest even as torch casts; NaN to the quiet NaN 0x7FC0);
 * host work only.  features NULL: the feature rows [0, n) are left as the
 * caller wrote them (a multi-threaded cast is faster than this single
 * thread); the padding rows are zeroed either way.
 * attgnn_serve_stage: the same into pinned host memory `pinned`, then one
 * hipMemcpyAsync of the buffer to `dev` on `stream`.  The caller keeps
 * `pinned` untouched until that copy has finished.
 */
int64_t attgnn_serve_stage_layout(int64_t n, int64_t nnz, int64_t feat, int window, int64_t *offsets);
int attgnn_serve_stage_host(const int64_t *indptr, const int32_t *indices, int64_t n, int64_t nnz,
                            const float *features, int64_t feat, int window, void *out, int64_t out_bytes);
int attgnn_serve_stage(const int64_t *indptr, const int32_t *indices, int64_t n, int64_t nnz,
                       const float *features, int64_t feat, int window, void *pinned, int64_t pinned_bytes,
                       void *dev, attgnn_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* ATTGNN_H */
